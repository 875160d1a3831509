// sdfx_server.cpp — the served form: a request loop over meshes kept resident
// on the device (no counterpart in the reference, which answers one call at a
// time on the caller's thread).
//
// Meshes are uploaded once (topology-order triangles for grid requests, the
// LBVH for point requests) and admitted while the resident bytes fit the HBM
// budget. Requests are queued by the callers' threads and served by worker
// threads, each with its own stream: a worker that comes free takes every
// queued request that fits the budget's free part (first come, first served)
// and runs them as one batch — the grid requests of each sign method through
// one batched grid run, the point requests of each query mode through one
// multi-mesh LBVH launch. Requests that arrive while batches run wait for the
// next free worker, and the batches of different workers overlap on the
// device. Every request's answer depends on its own mesh and inputs only, so
// it is bit-identical alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "sdfx_internal.hpp"

using namespace sdfx;
using Clock = std::chrono::steady_clock;

namespace {

// Admission estimates of device bytes (logical sizes; the caching allocator
// rounds blocks up to powers of two).
constexpr uint64_t MESH_BYTES_PER_TRI = 96 + 48 + 64;  // topology TriRecs, LBVH leaves, LBVH nodes
constexpr uint64_t GRID_BYTES_PER_CELL = 28;           // keys, flags, frontier items, ray parity bits
constexpr uint64_t GRID_BYTES_PER_TRI = 88;            // seed boxes, pair counts, owner table, 48-byte copies
constexpr uint64_t GRID_TILE_CELLS = 4096;             // each request's cells start on a tile
constexpr uint64_t POINT_BYTES_PER_QUERY = 56;         // sort keys, order, radix-sort buffers, deferred searches
constexpr uint64_t REQUEST_BYTES = 4096;               // tables and flags
// A batch's limits: the grid run's cell capacity and a cap on the table sizes.
constexpr uint64_t BATCH_MAX_CELLS = 150000000;
constexpr size_t BATCH_MAX_REQUESTS = 512;

struct ServerMesh {
    uint64_t nv = 0, nt = 0;
    uint64_t bytes = 0;
    DevBuf<TriRec> tris;  // topology order (grid requests, None point requests)
    Lbvh bvh;             // point requests
};

struct ServerRequest {
    uint64_t ticket = 0;
    bool grid = false;
    std::shared_ptr<ServerMesh> mesh;
    // point request
    const float* d_q = nullptr;
    uint64_t nq = 0;
    sdfx_acceleration_method method = SDFX_ACCEL_BVH;
    // grid request
    GridParams gp{};
    uint64_t cells = 0;
    bool zero_axis = false;
    sdfx_sign_method sign = SDFX_SIGN_RAYCAST;
    float* d_out = nullptr;
    uint64_t scratch = 0;
    Clock::time_point t_submit, t_done;
    bool done = false;
    sdfx_status status = SDFX_OK;
    std::string message;
};

uint64_t grid_scratch(uint64_t cells, uint64_t nt) {
    return (cells + GRID_TILE_CELLS) * GRID_BYTES_PER_CELL + nt * GRID_BYTES_PER_TRI + REQUEST_BYTES;
}

QueryMode point_mode(sdfx_acceleration_method m, sdfx_sign_method s) {
    switch (m) {
        case SDFX_ACCEL_BVH: return s == SDFX_SIGN_NORMAL ? QM_NORMAL_FOLD : QM_RAYCAST_VOTE;
        case SDFX_ACCEL_RTREE: return QM_NEAREST_SIGNED;
        default: return QM_NEAREST_VOTE;
    }
}

}  // namespace

struct sdfx_server {
    int device = 0;
    uint64_t budget = 0;
    uint64_t resident = 0, inflight = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<ServerRequest>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<ServerRequest>> tickets;
    std::unordered_map<uint32_t, std::shared_ptr<ServerMesh>> meshes;
    // Removed meshes that queued or running requests still hold: their bytes
    // stay in `resident` until the last such request lets go (release_retired).
    std::vector<std::shared_ptr<ServerMesh>> retired;
    std::vector<std::thread> workers;
    bool stop = false;
    uint64_t next_ticket = 1;
    uint32_t next_mesh = 1;
    sdfx_server_stats stats{};

    // The front request may start: it fits the free part of the budget, or
    // nothing else is in flight (so a request admitted at submission, which fits
    // the budget beside the resident meshes, always runs eventually).
    bool front_admissible() const {
        if (queue.empty()) return false;
        const uint64_t free_bytes = budget > resident ? budget - resident : 0;
        return inflight == 0 || inflight + queue.front()->scratch <= free_bytes;
    }

    // Takes the batch a free worker runs (called with mu held).
    std::vector<std::shared_ptr<ServerRequest>> take(uint64_t& bytes) {
        std::vector<std::shared_ptr<ServerRequest>> batch;
        const uint64_t free_bytes = budget > resident ? budget - resident : 0;
        uint64_t cells = 0;
        bytes = 0;
        while (!queue.empty() && batch.size() < BATCH_MAX_REQUESTS) {
            const auto& r = queue.front();
            const bool fits = inflight + bytes + r->scratch <= free_bytes;
            if (!batch.empty() && (!fits || cells + r->cells > BATCH_MAX_CELLS)) break;
            if (batch.empty() && !fits && inflight > 0) break;
            cells += r->cells;
            bytes += r->scratch;
            batch.push_back(r);
            queue.pop_front();
        }
        inflight += bytes;
        stats.inflight_bytes = inflight;
        stats.peak_inflight_bytes = std::max(stats.peak_inflight_bytes, inflight);
        return batch;
    }

    void worker_main();
    // Per worker: the grid stream, the point stream and two events ordering them.
    struct WorkerStreams {
        hipStream_t grid = nullptr, points = nullptr;
        hipEvent_t flags_ready = nullptr, points_done = nullptr;
    };
    void run_batch(std::vector<std::shared_ptr<ServerRequest>>& batch, const WorkerStreams& ws);
    // Drops the retired meshes no request holds any more and returns their
    // bytes to the budget (with `mu` held).
    void release_retired() {
        for (size_t k = 0; k < retired.size();) {
            if (retired[k].use_count() == 1) {
                resident -= retired[k]->bytes;
                stats.resident_bytes = resident;
                retired[k] = std::move(retired.back());
                retired.pop_back();
            } else {
                ++k;
            }
        }
    }
};

// A batch's point requests (None and LBVH) run on the worker's point stream
// while the host drives the grid rounds on its grid stream: the grid rounds are
// bound by launch and memory latency and leave most of the GPU to the point
// traversals. The two streams meet again before the flags are read.
void sdfx_server::run_batch(std::vector<std::shared_ptr<ServerRequest>>& batch, const WorkerStreams& ws) {
    const size_t n = batch.size();
    const hipStream_t s = ws.grid;
    std::vector<DevFlags> hf(n);
    {
        Arena A(s);
        Arena B(ws.points);  // destroyed first: synchronises the point stream before A's memory is released
        DevFlags* flags = A.alloc<DevFlags>(n);
        fill_u32(reinterpret_cast<uint32_t*>(flags), n * (sizeof(DevFlags) / sizeof(uint32_t)), 0u, s);
        SDFX_HIP(hipEventRecord(ws.flags_ready, s));
        SDFX_HIP(hipStreamWaitEvent(ws.points, ws.flags_ready, 0));
        std::vector<GridBatchItem> grid[2];
        uint64_t tri_off[2] = {0, 0};
        std::vector<PointBatchItem> points[4];
        for (size_t i = 0; i < n; ++i) {
            ServerRequest& r = *batch[i];
            const ServerMesh& m = *r.mesh;
            if (r.grid) {
                if (r.zero_axis) continue;  // settled below: nothing to compute
                const int g = r.sign == SDFX_SIGN_NORMAL ? 1 : 0;
                grid[g].push_back(GridBatchItem{m.tris.get(), tri_off[g], m.nt, r.gp, r.d_out, (uint32_t)i});
                tri_off[g] += m.nt;
            } else if (r.method == SDFX_ACCEL_NONE) {
                none_query(m.tris.get(), m.nt, r.d_q, r.nq, r.sign, r.d_out, flags + i, B);
            } else if (m.nt > 0 && r.nq > 0) {
                points[point_mode(r.method, r.sign)].push_back(PointBatchItem{&m.bvh, r.d_q, r.nq, r.d_out,
                                                                              (uint32_t)i});
            }
        }
        for (int q = 0; q < 4; ++q)  // queued on the point stream without blocking the host
            if (!points[q].empty()) lbvh_query_batch(points[q], (QueryMode)q, flags, B);
        SDFX_HIP(hipEventRecord(ws.points_done, ws.points));
        for (int g = 0; g < 2; ++g)
            if (!grid[g].empty()) grid_sdf_batch(grid[g], g ? SDFX_SIGN_NORMAL : SDFX_SIGN_RAYCAST, flags, A, nullptr);
        SDFX_HIP(hipStreamWaitEvent(s, ws.points_done, 0));
        SDFX_HIP(hipMemcpyAsync(hf.data(), flags, n * sizeof(DevFlags), hipMemcpyDeviceToHost, s));
        SDFX_HIP(hipStreamSynchronize(s));
    }
    for (size_t i = 0; i < n; ++i) {
        ServerRequest& r = *batch[i];
        const DevFlags& f = hf[i];
        if (f.internal) {
            r.status = SDFX_ERR_DEVICE;
            r.message = "device consistency check failed (traversal stack or work decode)";
        } else if (r.grid && r.zero_axis && r.mesh->nt > 0) {
            r.status = SDFX_ERR_INVALID_GRID;  // grid.rs:311-315
            r.message = "assertion failed: min <= max";
        } else if (f.nan_distance) {
            r.status = SDFX_ERR_NAN_DISTANCE;
            r.message = "NaN distance";
        }
    }
}

void sdfx_server::worker_main() {
    WorkerStreams ws;
    bool ready = false;
    try {
        SDFX_HIP(hipSetDevice(device));
        SDFX_HIP(hipStreamCreateWithFlags(&ws.grid, hipStreamNonBlocking));
        SDFX_HIP(hipStreamCreateWithFlags(&ws.points, hipStreamNonBlocking));
        SDFX_HIP(hipEventCreateWithFlags(&ws.flags_ready, hipEventDisableTiming));
        SDFX_HIP(hipEventCreateWithFlags(&ws.points_done, hipEventDisableTiming));
        ready = true;
    } catch (const Error&) {
        ready = false;
    }
    for (;;) {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return (stop && queue.empty()) || front_admissible(); });
        if (queue.empty()) break;  // stopping
        uint64_t bytes = 0;
        std::vector<std::shared_ptr<ServerRequest>> batch = take(bytes);
        lk.unlock();
        sdfx_status st = SDFX_OK;
        std::string msg;
        if (!ready) {
            st = SDFX_ERR_DEVICE;
            msg = "server worker could not create its streams";
        } else {
            try {
                run_batch(batch, ws);
            } catch (const Error& e) {
                st = e.status;
                msg = e.what();
            } catch (const std::exception& e) {
                st = SDFX_ERR_DEVICE;
                msg = e.what();
            }
        }
        const Clock::time_point t = Clock::now();
        lk.lock();
        for (auto& r : batch) {
            if (st != SDFX_OK) {
                r->status = st;
                r->message = msg;
            }
            r->t_done = t;
            r->done = true;
            r->mesh.reset();  // a removed mesh is freed once no request holds it
        }
        release_retired();
        inflight -= bytes;
        stats.inflight_bytes = inflight;
        stats.requests_done += batch.size();
        stats.batches += 1;
        stats.max_batch_requests = std::max<uint64_t>(stats.max_batch_requests, batch.size());
        lk.unlock();
        cv_done.notify_all();
        cv_work.notify_all();
    }
    if (ws.flags_ready) (void)hipEventDestroy(ws.flags_ready);
    if (ws.points_done) (void)hipEventDestroy(ws.points_done);
    if (ws.points) (void)hipStreamDestroy(ws.points);
    if (ws.grid) (void)hipStreamDestroy(ws.grid);
}

namespace {

template <typename F>
sdfx_status server_guarded(F&& f) {
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
    }
}

[[noreturn]] void server_fail(sdfx_status s, const std::string& msg) { throw Error(s, msg); }

std::shared_ptr<ServerMesh> find_mesh(sdfx_server* sv, uint32_t id) {
    auto it = sv->meshes.find(id);
    if (it == sv->meshes.end()) server_fail(SDFX_ERR_INVALID_ARGUMENT, "unknown mesh id " + std::to_string(id));
    return it->second;
}

// Queues r (or completes it at once when it has no work) and returns its ticket.
uint64_t enqueue(sdfx_server* sv, std::shared_ptr<ServerRequest> r, bool immediate) {
    std::lock_guard<std::mutex> lk(sv->mu);
    r->ticket = sv->next_ticket++;
    r->t_submit = Clock::now();
    sv->tickets[r->ticket] = r;
    if (immediate) {
        r->t_done = r->t_submit;
        r->done = true;
    } else {
        sv->queue.push_back(r);
        sv->cv_work.notify_one();
    }
    return r->ticket;
}

void check_fits(sdfx_server* sv, uint64_t scratch) {
    std::lock_guard<std::mutex> lk(sv->mu);
    const uint64_t free_bytes = sv->budget > sv->resident ? sv->budget - sv->resident : 0;
    if (scratch > free_bytes)
        server_fail(SDFX_ERR_OVER_BUDGET, "request needs " + std::to_string(scratch) + " bytes of scratch; " +
                                              std::to_string(free_bytes) + " of the budget are not held by meshes");
}

}  // namespace

extern "C" {

sdfx_status sdfx_server_create(uint64_t hbm_budget_bytes, uint32_t workers, sdfx_server** server_out) {
    return server_guarded([&]() -> sdfx_status {
        if (!server_out) server_fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: server_out");
        *server_out = nullptr;
        if (workers == 0 || workers > 8) server_fail(SDFX_ERR_INVALID_ARGUMENT, "workers must be in [1, 8]");
        std::unique_ptr<sdfx_server> sv(new sdfx_server());
        sv->device = current_device();
        sv->budget = hbm_budget_bytes;
        for (uint32_t w = 0; w < workers; ++w) sv->workers.emplace_back([p = sv.get()] { p->worker_main(); });
        *server_out = sv.release();
        return SDFX_OK;
    });
}

void sdfx_server_destroy(sdfx_server* sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    for (auto& t : sv->workers) t.join();
    (void)hipSetDevice(sv->device);
    delete sv;
}

}  // extern "C"

namespace {

// Adds meshes[0, count): validates each, admits them in order while they fit
// the budget beside the resident meshes (then waits for the requests in flight
// to leave room), gathers their triangles and builds all their LBVHs in one
// batched build. Fills each mesh's id and status; message[i] holds the
// reference's message for a failed one.
void add_meshes_core(sdfx_server* sv, sdfx_server_mesh* meshes, uint64_t count, std::vector<std::string>& message) {
    message.assign(count, std::string());
    std::vector<uint64_t> nt(count, 0), bytes(count, 0);
    for (uint64_t i = 0; i < count; ++i) {
        sdfx_server_mesh& d = meshes[i];
        d.id = 0;
        d.status = SDFX_OK;
        if (!d.vertices && d.vertex_count) {
            d.status = SDFX_ERR_INVALID_ARGUMENT;
            message[i] = "null pointer: vertices";
        } else if (d.topology != SDFX_TRIANGLE_LIST && d.topology != SDFX_TRIANGLE_STRIP) {
            d.status = SDFX_ERR_INVALID_ARGUMENT;
            message[i] = "unknown topology";
        } else {
            const sdfx_status ts =
                topology_triangle_count(d.vertex_count, d.indices != nullptr, d.index_count, d.topology, &nt[i]);
            if (ts == SDFX_ERR_INDEX_COUNT)
                message[i] = "TriangleList indices length (" + std::to_string(d.index_count) + ") must be divisible by 3";
            else if (ts == SDFX_ERR_VERTEX_COUNT)
                message[i] = "TriangleList vertices length (" + std::to_string(d.vertex_count) +
                             ") must be divisible by 3 when no indices are provided";
            else if (ts != SDFX_OK)
                message[i] = "invalid topology";
            d.status = ts;
        }
        bytes[i] = nt[i] * MESH_BYTES_PER_TRI + REQUEST_BYTES;
    }
    uint64_t reserved = 0;
    {
        // admitted while they fit beside the resident meshes; then the requests
        // in flight (which always drain) must leave room
        std::unique_lock<std::mutex> lk(sv->mu);
        for (uint64_t i = 0; i < count; ++i) {
            if (meshes[i].status != SDFX_OK) continue;
            if (sv->resident + reserved + bytes[i] > sv->budget) {
                meshes[i].status = SDFX_ERR_OVER_BUDGET;
                message[i] = "mesh needs " + std::to_string(bytes[i]) + " bytes; " +
                             std::to_string(sv->budget - std::min(sv->budget, sv->resident + reserved)) +
                             " of the budget are not held by meshes";
                continue;
            }
            reserved += bytes[i];
        }
        sv->cv_done.wait(lk, [&] { return sv->resident + sv->inflight + reserved <= sv->budget; });
        sv->resident += reserved;
        sv->stats.resident_bytes = sv->resident;
    }
    auto release = [&](uint64_t b) {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->resident -= b;
        sv->stats.resident_bytes = sv->resident;
    };
    std::vector<std::shared_ptr<ServerMesh>> built(count);
    try {
        SDFX_HIP(hipSetDevice(sv->device));
        hipStream_t s = thread_stream();
        std::vector<DevFlags> hf(count ? count : 1);
        {
            Arena A(s);
            DevFlags* flags = A.alloc<DevFlags>(count ? count : 1);
            fill_u32(reinterpret_cast<uint32_t*>(flags), (count ? count : 1) * (sizeof(DevFlags) / sizeof(uint32_t)),
                     0u, s);
            std::vector<LbvhBuildJob> jobs;
            std::vector<GatherJob> gathers;
            for (uint64_t i = 0; i < count; ++i) {
                if (meshes[i].status != SDFX_OK) continue;
                const sdfx_server_mesh& d = meshes[i];
                auto m = std::make_shared<ServerMesh>();
                m->nv = d.vertex_count;
                m->nt = nt[i];
                m->bytes = bytes[i];
                m->tris.reset(nt[i]);
                gathers.push_back(GatherJob{d.vertices, d.vertex_count, d.indices, d.topology, nt[i], m->tris.get(),
                                            flags + i});
                if (nt[i] > 0) jobs.push_back(LbvhBuildJob{m->tris.get(), nt[i], &m->bvh});
                built[i] = std::move(m);
            }
            gather_triangles_batch(gathers, A);
            lbvh_build_batch(jobs, A);
            SDFX_HIP(hipMemcpyAsync(hf.data(), flags, count * sizeof(DevFlags), hipMemcpyDeviceToHost, s));
            SDFX_HIP(hipStreamSynchronize(s));
        }
        for (uint64_t i = 0; i < count; ++i) {
            if (!built[i] || !hf[i].index_out_of_range) continue;
            meshes[i].status = SDFX_ERR_INDEX_OUT_OF_RANGE;
            message[i] = "index out of bounds: the len is " + std::to_string(meshes[i].vertex_count) +
                         " but a triangle index is >= " + std::to_string(meshes[i].vertex_count);
            release(bytes[i]);
            built[i].reset();
        }
    } catch (...) {
        uint64_t held = 0;
        for (uint64_t i = 0; i < count; ++i)
            if (built[i] || meshes[i].status == SDFX_OK) held += bytes[i];
        release(held);
        throw;
    }
    std::lock_guard<std::mutex> lk(sv->mu);
    for (uint64_t i = 0; i < count; ++i) {
        if (!built[i]) continue;
        const uint32_t id = sv->next_mesh++;
        sv->meshes[id] = std::move(built[i]);
        meshes[i].id = id;
    }
    sv->stats.meshes = sv->meshes.size();
}

}  // namespace

extern "C" {

sdfx_status sdfx_server_add_meshes(sdfx_server* sv, sdfx_server_mesh* meshes, uint64_t count) {
    return server_guarded([&]() -> sdfx_status {
        if (!sv || (!meshes && count)) server_fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: server or meshes");
        std::vector<std::string> message;
        add_meshes_core(sv, meshes, count, message);
        for (uint64_t i = 0; i < count; ++i)
            if (meshes[i].status != SDFX_OK) {
                set_last_error("mesh " + std::to_string(i) + ": " + message[i]);
                return meshes[i].status;
            }
        return SDFX_OK;
    });
}

sdfx_status sdfx_server_add_mesh(sdfx_server* sv, const float* d_vertices, uint64_t vertex_count,
                                 const uint32_t* d_indices, uint64_t index_count, sdfx_topology topology,
                                 uint32_t* mesh_id_out) {
    return server_guarded([&]() -> sdfx_status {
        if (!sv || !mesh_id_out) server_fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: server or mesh_id_out");
        sdfx_server_mesh d{d_vertices, vertex_count, d_indices, index_count, topology, 0u, SDFX_OK};
        std::vector<std::string> message;
        add_meshes_core(sv, &d, 1, message);
        if (d.status != SDFX_OK) server_fail(d.status, message[0]);
        *mesh_id_out = d.id;
        return SDFX_OK;
    });
}

sdfx_status sdfx_server_remove_mesh(sdfx_server* sv, uint32_t mesh_id) {
    return server_guarded([&]() -> sdfx_status {
        if (!sv) server_fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: server");
        std::lock_guard<std::mutex> lk(sv->mu);
        auto m = find_mesh(sv, mesh_id);
        sv->meshes.erase(mesh_id);
        // queued and running requests hold the mesh until they finish: its
        // bytes leave the budget when the last of them does
        sv->retired.push_back(std::move(m));
        sv->release_retired();
        sv->stats.meshes = sv->meshes.size();
        sv->cv_work.notify_all();
        return SDFX_OK;
    });
}

sdfx_status sdfx_server_submit_points(sdfx_server* sv, uint32_t mesh_id, const float* d_query_points,
                                      uint64_t query_count, sdfx_acceleration_method method, sdfx_sign_method sign,
                                      float* d_out, uint64_t* ticket_out) {
    return server_guarded([&]() -> sdfx_status {
        if (!sv || !ticket_out) server_fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: server or ticket_out");
        if ((!d_query_points || !d_out) && query_count)
            server_fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: query_points or out");
        if (method < SDFX_ACCEL_NONE || method > SDFX_ACCEL_RTREE_BVH)
            server_fail(SDFX_ERR_INVALID_ARGUMENT, "unknown acceleration method");
        if (sign != SDFX_SIGN_RAYCAST && sign != SDFX_SIGN_NORMAL)
            server_fail(SDFX_ERR_INVALID_ARGUMENT, "unknown sign method");
        auto r = std::make_shared<ServerRequest>();
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            r->mesh = find_mesh(sv, mesh_id);
        }
        r->d_q = d_query_points;
        r->nq = query_count;
        r->method = method;
        // Rtree always signs by normal, RtreeBvh by raycast (lib.rs:272-286)
        r->sign = method == SDFX_ACCEL_RTREE ? SDFX_SIGN_NORMAL
                                             : (method == SDFX_ACCEL_RTREE_BVH ? SDFX_SIGN_RAYCAST : sign);
        r->d_out = d_out;
        r->scratch = query_count * POINT_BYTES_PER_QUERY + REQUEST_BYTES;
        bool immediate = query_count == 0;
        if (!immediate && method != SDFX_ACCEL_NONE && r->mesh->nt == 0) {
            // generic/rtree_bvh.rs:96-98 returns an empty Vec; Rtree and Bvh
            // unwrap a missing nearest triangle
            immediate = true;
            if (method == SDFX_ACCEL_RTREE_BVH) {
                r->status = SDFX_EMPTY_RESULT;
            } else {
                r->status = SDFX_ERR_EMPTY_MESH;
                r->message = "called `Option::unwrap()` on a `None` value";
            }
        }
        if (!immediate) check_fits(sv, r->scratch);
        *ticket_out = enqueue(sv, r, immediate);
        return SDFX_OK;
    });
}

sdfx_status sdfx_server_submit_grid(sdfx_server* sv, uint32_t mesh_id, const float first_cell[3],
                                    const float cell_size[3], const uint64_t cell_count[3], sdfx_sign_method sign,
                                    float* d_out, uint64_t* ticket_out) {
    return server_guarded([&]() -> sdfx_status {
        if (!sv || !ticket_out) server_fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: server or ticket_out");
        if (!first_cell || !cell_size || !cell_count) server_fail(SDFX_ERR_INVALID_ARGUMENT, "null grid description");
        if (sign != SDFX_SIGN_RAYCAST && sign != SDFX_SIGN_NORMAL)
            server_fail(SDFX_ERR_INVALID_ARGUMENT, "unknown sign method");
        auto r = std::make_shared<ServerRequest>();
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            r->mesh = find_mesh(sv, mesh_id);
        }
        r->grid = true;
        r->sign = sign;
        r->d_out = d_out;
        for (int k = 0; k < 3; ++k) {
            r->gp.first[k] = first_cell[k];
            r->gp.size[k] = cell_size[k];
            r->gp.count[k] = cell_count[k];
        }
        r->zero_axis = cell_count[0] == 0 || cell_count[1] == 0 || cell_count[2] == 0;
        if (!r->zero_axis) {
            if (cell_count[0] > (1u << 21) || cell_count[1] > (1u << 21) || cell_count[2] > (1u << 22))
                server_fail(SDFX_ERR_INVALID_ARGUMENT, "grid axes are limited to 2^21 (x, y) and 2^22 (z) cells");
            r->cells = cell_count[0] * cell_count[1] * cell_count[2];
            if (r->cells > BATCH_MAX_CELLS)
                server_fail(SDFX_ERR_INVALID_ARGUMENT, "served grids are limited to 150,000,000 cells");
            if (!d_out) server_fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: out");
        }
        r->scratch = grid_scratch(r->cells, r->mesh->nt);
        const bool immediate = r->zero_axis;
        if (immediate && r->mesh->nt > 0) {
            r->status = SDFX_ERR_INVALID_GRID;  // grid.rs:311-315
            r->message = "assertion failed: min <= max";
        }
        if (!immediate) check_fits(sv, r->scratch);
        *ticket_out = enqueue(sv, r, immediate);
        return SDFX_OK;
    });
}

sdfx_status sdfx_server_wait(sdfx_server* sv, uint64_t ticket, double* latency_ms) {
    return server_guarded([&]() -> sdfx_status {
        if (!sv) server_fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: server");
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->tickets.find(ticket);
        if (it == sv->tickets.end()) server_fail(SDFX_ERR_INVALID_ARGUMENT, "unknown ticket " + std::to_string(ticket));
        std::shared_ptr<ServerRequest> r = it->second;
        sv->cv_done.wait(lk, [&] { return r->done; });
        sv->tickets.erase(ticket);
        if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
        if (r->status < 0) set_last_error(r->message);
        return r->status;
    });
}

sdfx_status sdfx_server_cancel(sdfx_server* sv, uint64_t ticket, int* cancelled_out) {
    return server_guarded([&]() -> sdfx_status {
        if (!sv || !cancelled_out) server_fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: server or cancelled_out");
        *cancelled_out = 0;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            auto it = sv->tickets.find(ticket);
            if (it == sv->tickets.end())
                server_fail(SDFX_ERR_INVALID_ARGUMENT, "unknown ticket " + std::to_string(ticket));
            const std::shared_ptr<ServerRequest> r = it->second;
            auto q = std::find(sv->queue.begin(), sv->queue.end(), r);
            if (q == sv->queue.end()) return SDFX_OK;  // running or done
            sv->queue.erase(q);
            r->mesh.reset();
            sv->release_retired();
            r->status = SDFX_ERR_CANCELLED;
            r->message = "request cancelled";
            r->t_done = Clock::now();
            r->done = true;
            *cancelled_out = 1;
        }
        sv->cv_done.notify_all();
        sv->cv_work.notify_all();  // the queue's front may have changed
        return SDFX_OK;
    });
}

sdfx_status sdfx_server_poll(sdfx_server* sv, uint64_t ticket, int* done_out) {
    return server_guarded([&]() -> sdfx_status {
        if (!sv || !done_out) server_fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: server or done_out");
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->tickets.find(ticket);
        if (it == sv->tickets.end()) server_fail(SDFX_ERR_INVALID_ARGUMENT, "unknown ticket " + std::to_string(ticket));
        *done_out = it->second->done ? 1 : 0;
        return SDFX_OK;
    });
}

sdfx_status sdfx_server_get_stats(sdfx_server* sv, sdfx_server_stats* out) {
    return server_guarded([&]() -> sdfx_status {
        if (!sv || !out) server_fail(SDFX_ERR_INVALID_ARGUMENT, "null pointer: server or out");
        std::lock_guard<std::mutex> lk(sv->mu);
        *out = sv->stats;
        out->queued = sv->queue.size();
        return SDFX_OK;
    });
}

}  // extern "C"
