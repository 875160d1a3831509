/*
 * sdfx.h — C ABI of libsdfx, the MI355X (gfx950) implementation of
 * mesh_to_sdf's hot path.
 *
 * Every entry point below stands in for one function of the reference crate's
 * public Rust API; the comment above each cites the interface it replaces
 * (paths relative to mesh_to_sdf/src/ in ZiplineTeam/mesh_to_sdf). The
 * signatures use plain pointers and sizes only, so a Rust `-sys` crate,
 * ctypes or cgo can bind them without HIP headers.
 *
 * Data layout: vertices and query points are packed f32 triples (x, y, z),
 * i.e. `[[f32; 3]]`; indices are u32. Grid output is f32 per cell with z the
 * fastest axis (grid.rs:276-278).
 *
 * Host-pointer entry points copy their inputs to the device, run, copy the
 * result back and return when it is in `out` (the reference is synchronous).
 * `_device` entry points take device pointers and a hipStream_t (passed as
 * void*), enqueue everything on that stream and synchronise it before
 * returning the status.
 *
 * Errors: the reference panics; here every panic becomes a negative status
 * and a thread-local message (sdfx_last_error). Conditions are checked in the
 * order the reference would hit them.
 *
 * Thread safety: all entry points are re-entrant. Each calling thread gets its
 * own stream for host-pointer calls; sdfx_mesh handles are immutable after
 * build and may be queried from several threads.
 */
#ifndef SDFX_H
#define SDFX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDFX_VERSION 1

typedef enum sdfx_status {
    SDFX_OK = 0,
    /* Not an error: the reference returns an empty Vec / None here
     * (generate/generic/rtree_bvh.rs:96-98, 124-126). *out_count is 0. */
    SDFX_EMPTY_RESULT = 1,
    /* Not an error: check_mesh_triangle_list returned Err (the message says why). */
    SDFX_MESH_INVALID = 2,
    /* lib.rs:375 "Query point {..} contains non-finite values" */
    SDFX_ERR_NON_FINITE_QUERY = -1,
    /* lib.rs:203 "TriangleList indices length ({}) must be divisible by 3" */
    SDFX_ERR_INDEX_COUNT = -2,
    /* lib.rs:207 "TriangleList vertices length ({}) must be divisible by 3 ..." */
    SDFX_ERR_VERTEX_COUNT = -3,
    /* slice index out of bounds when a triangle reads vertices[i] */
    SDFX_ERR_INDEX_OUT_OF_RANGE = -4,
    /* generic/rtree.rs:148 nearest_neighbor().unwrap() on an empty tree */
    SDFX_ERR_EMPTY_MESH = -5,
    /* lib.rs:330 "NaN distance" */
    SDFX_ERR_NAN_DISTANCE = -6,
    SDFX_ERR_INVALID_ARGUMENT = -7,
    SDFX_ERR_DEVICE = -8,
    SDFX_ERR_OUT_OF_MEMORY = -9,
    /* grid.rs:311-315 clamp(0, n - 1) panics on a zero cell count */
    SDFX_ERR_INVALID_GRID = -10,
    /* sdfx_server_*: the mesh or request does not fit the server's HBM budget */
    SDFX_ERR_OVER_BUDGET = -11,
    /* sdfx_server_*: the request was cancelled before it ran */
    SDFX_ERR_CANCELLED = -12
} sdfx_status;

/* lib.rs:156-171 Topology::{TriangleList, TriangleStrip}; indices == NULL is `None` */
typedef enum sdfx_topology { SDFX_TRIANGLE_LIST = 0, SDFX_TRIANGLE_STRIP = 1 } sdfx_topology;

/* lib.rs:252-263 SignMethod (Raycast is the default) */
typedef enum sdfx_sign_method { SDFX_SIGN_RAYCAST = 0, SDFX_SIGN_NORMAL = 1 } sdfx_sign_method;

/* lib.rs:272-286 AccelerationMethod; None and Bvh carry a sign method, Rtree
 * always uses Normal and RtreeBvh always uses Raycast (the sign argument is
 * ignored for those two). */
typedef enum sdfx_acceleration_method {
    SDFX_ACCEL_NONE = 0,
    SDFX_ACCEL_BVH = 1,
    SDFX_ACCEL_RTREE = 2,
    SDFX_ACCEL_RTREE_BVH = 3
} sdfx_acceleration_method;

/* Opaque resident-mesh handle: lib.rs:303-312 SdfAccelerationMesh. Owns the
 * device copy of the mesh and its device LBVH until sdfx_free_sdf_acceleration. */
typedef struct sdfx_mesh sdfx_mesh;

/* Message of the last failing call on this thread ("" if none). */
const char* sdfx_last_error(void);

/* SDFX_VERSION of the loaded library. */
int sdfx_version(void);

/* Replaces lib.rs:364-391 generate_sdf<V, I>(vertices, indices, query_points,
 * acceleration_method) -> Vec<f32>. `out` holds query_count floats;
 * *out_count receives the number written (query_count, or 0 with
 * SDFX_EMPTY_RESULT). */
sdfx_status sdfx_generate_sdf(const float* vertices, uint64_t vertex_count, const uint32_t* indices,
                              uint64_t index_count, sdfx_topology topology, const float* query_points,
                              uint64_t query_count, sdfx_acceleration_method method, sdfx_sign_method sign,
                              float* out, uint64_t* out_count);

/* Device-pointer form of sdfx_generate_sdf (same semantics; `stream` is a
 * hipStream_t or NULL for the null stream). */
sdfx_status sdfx_generate_sdf_device(const float* d_vertices, uint64_t vertex_count, const uint32_t* d_indices,
                                     uint64_t index_count, sdfx_topology topology, const float* d_query_points,
                                     uint64_t query_count, sdfx_acceleration_method method,
                                     sdfx_sign_method sign, float* d_out, uint64_t* out_count, void* stream);

/* Replaces generate/grid.rs:265-273 generate_grid_sdf<V, I>(vertices, indices,
 * grid, sign_method) -> Vec<f32>, with the grid given as Grid::new(first_cell,
 * cell_size, cell_count) (grid.rs:197-203). `out` holds
 * cell_count[0]*cell_count[1]*cell_count[2] floats, z fastest. */
sdfx_status sdfx_generate_grid_sdf(const float* vertices, uint64_t vertex_count, const uint32_t* indices,
                                   uint64_t index_count, sdfx_topology topology, const float first_cell[3],
                                   const float cell_size[3], const uint64_t cell_count[3], sdfx_sign_method sign,
                                   float* out);

/* Device-pointer form of sdfx_generate_grid_sdf. */
sdfx_status sdfx_generate_grid_sdf_device(const float* d_vertices, uint64_t vertex_count,
                                          const uint32_t* d_indices, uint64_t index_count, sdfx_topology topology,
                                          const float first_cell[3], const float cell_size[3],
                                          const uint64_t cell_count[3], sdfx_sign_method sign, float* d_out,
                                          void* stream);

/* No counterpart in the reference: a schedule-free generate_grid_sdf for
 * callers that need each cell's distance exact rather than the flood fill's
 * approximation, whose error in the reference depends on its thread count.
 * Same arguments, errors and sign as sdfx_generate_grid_sdf; each cell's
 * magnitude is the FP32 minimum over all triangles of geo.rs's distance from
 * the cell centre (Raycast, signs from the same grid raycasts as
 * generate/grid.rs:568-684), or the compare_distances fold over all triangles
 * of geo.rs's signed distance (Normal, as Bvh(Normal) answers a query point,
 * generic/bvh.rs:117-129). */
sdfx_status sdfx_generate_grid_sdf_exact(const float* vertices, uint64_t vertex_count, const uint32_t* indices,
                                         uint64_t index_count, sdfx_topology topology, const float first_cell[3],
                                         const float cell_size[3], const uint64_t cell_count[3],
                                         sdfx_sign_method sign, float* out);

/* Device-pointer form of sdfx_generate_grid_sdf_exact. */
sdfx_status sdfx_generate_grid_sdf_exact_device(const float* d_vertices, uint64_t vertex_count,
                                                const uint32_t* d_indices, uint64_t index_count,
                                                sdfx_topology topology, const float first_cell[3],
                                                const float cell_size[3], const uint64_t cell_count[3],
                                                sdfx_sign_method sign, float* d_out, void* stream);

/* One request of a batched generate_grid_sdf (the fields of
 * sdfx_generate_grid_sdf's arguments); `status` is set by the call. */
typedef struct sdfx_grid_request {
    const float* vertices;
    uint64_t vertex_count;
    const uint32_t* indices; /* NULL: no index buffer */
    uint64_t index_count;
    sdfx_topology topology;
    float first_cell[3];
    float cell_size[3];
    uint64_t cell_count[3];
    float* out;
    sdfx_status status;
} sdfx_grid_request;

/* Batched generate_grid_sdf (no counterpart in the reference, which takes one
 * mesh per call; for callers with many meshes): every request, all with one
 * sign method, runs through the same device launches, and each output is
 * bit-identical to sdfx_generate_grid_sdf on that request alone. Returns
 * SDFX_OK when every request succeeded, else the first failing request's status
 * (each request's own status is in its `status`; sdfx_last_error names the
 * request). Pointers are host pointers; the _device form takes device pointers
 * and a stream. */
sdfx_status sdfx_generate_grid_sdf_batch(sdfx_grid_request* requests, uint64_t request_count,
                                         sdfx_sign_method sign);
sdfx_status sdfx_generate_grid_sdf_batch_device(sdfx_grid_request* requests, uint64_t request_count,
                                                sdfx_sign_method sign, void* stream);

/* Replaces lib.rs:410-433 build_sdf_acceleration(vertices, indices,
 * acceleration_method) -> Option<SdfAccelerationMesh>. *mesh_out is NULL with
 * SDFX_EMPTY_RESULT where the reference returns None (RtreeBvh on an empty
 * mesh). */
sdfx_status sdfx_build_sdf_acceleration(const float* vertices, uint64_t vertex_count, const uint32_t* indices,
                                        uint64_t index_count, sdfx_topology topology,
                                        sdfx_acceleration_method method, sdfx_sign_method sign,
                                        sdfx_mesh** mesh_out);

/* Replaces lib.rs:463-485 query_sdf(&SdfAccelerationMesh, query_points) -> Vec<f32>.
 * Like the reference, no finiteness check is made on the queries. */
sdfx_status sdfx_query_sdf(const sdfx_mesh* mesh, const float* query_points, uint64_t query_count, float* out,
                           uint64_t* out_count);

/* sdfx_query_sdf on device memory with no allocation at all (no counterpart
 * in the reference: for callers that cannot allocate device memory, such as
 * code under HIP graph capture or serving loops that forbid it). The call's
 * temporaries come from a fixed 16 MiB scratch array that each GPU holds
 * while the library is loaded; at most 262,144 queries per call (more:
 * SDFX_ERR_INVALID_ARGUMENT). Points and outputs are device pointers, work is
 * enqueued on `stream` (a hipStream_t, NULL for the null stream), and the call
 * returns when it is done. Calls on one device share that scratch, so the
 * library runs them one after another. While `stream` is being captured
 * (hipStreamBeginCapture), the call records its work into the graph and
 * returns SDFX_OK at once: the device-side checks (a NaN distance, which the
 * reference panics on) are then not reported, and every replay uses the same
 * scratch, so replays that overlap with each other or with other calls on one
 * GPU race. */
sdfx_status sdfx_query_sdf_noalloc(const sdfx_mesh* mesh, const float* d_query_points, uint64_t query_count,
                                   float* d_out, uint64_t* out_count, void* stream);

/* One request of sdfx_query_sdf_batch_device: a resident mesh, its query
 * points and the output (device pointers); `status` is set by the call. */
typedef struct sdfx_point_request {
    const sdfx_mesh* mesh;
    const float* query_points;
    uint64_t query_count;
    float* out;
    sdfx_status status;
} sdfx_point_request;

/* query_sdf (lib.rs:463-485) for many resident meshes with uneven query
 * counts in one call (no counterpart in the reference, which answers one mesh
 * per call; for callers with many meshes): the requests of each query mode
 * share the same LBVH launches (queries sorted by request, then Morton code),
 * and each output is bit-identical to sdfx_query_sdf on that request alone.
 * All meshes must live on one device. Returns SDFX_OK when every request
 * succeeded, else the first failing request's status (each request's own
 * status is in its `status`; sdfx_last_error names the request). Work is
 * enqueued on `stream` and the call returns when it is done. */
sdfx_status sdfx_query_sdf_batch_device(sdfx_point_request* requests, uint64_t request_count, void* stream);

/* Replaces generic/rtree_bvh.rs:217-276 query_vector_to_closest_point_rtree_bvh
 * (re-exported at lib.rs:150). `mesh` must have been built with
 * SDFX_ACCEL_RTREE_BVH. vectors_out holds 3*query_count floats (closest - query),
 * inside_out query_count bytes (1 = inside by the 3-ray vote). */
sdfx_status sdfx_query_vector_to_closest_point(const sdfx_mesh* mesh, const float* query_points,
                                               uint64_t query_count, float* vectors_out, uint8_t* inside_out);

/* Drops an SdfAccelerationMesh (lib.rs:303-312 is Clone + Drop). NULL is allowed. */
void sdfx_free_sdf_acceleration(sdfx_mesh* mesh);

/* Replaces generate/mesh_check.rs:24-81 check_mesh_triangle_list(vertices,
 * faces) -> Result<(), String> (re-exported at lib.rs:149) for u32 faces.
 * Returns SDFX_OK for Ok(()) and SDFX_MESH_INVALID for Err, with the reference's
 * message in `message` (NUL-terminated, truncated to message_capacity). The
 * optional kind_out gets 0 (Ok), 1 (face count not a multiple of 3),
 * 2 (duplicate vertex), 3 (duplicate face), 4 (duplicate edge) or 5 (edge
 * without its reverse); where_out the vertex index, face index or edge index
 * 3 * face + k of edge (face[k], face[(k + 1) % 3]). Runs on the device. */
sdfx_status sdfx_check_mesh_triangle_list(const float* vertices, uint64_t vertex_count, const uint32_t* faces,
                                          uint64_t face_index_count, int* kind_out, uint64_t* where_out,
                                          char* message, uint64_t message_capacity);

/* ---- Tuning (no counterpart in the reference). Process-wide knobs, none of
 * which changes a result; returns 0 on success, -1 for an unknown key or a
 * value out of range.
 *   "exhaustive_max_pairs": queries x triangles at or below which Bvh, Rtree
 *   and RtreeBvh evaluate every triangle instead of building the LBVH (results
 *   are identical either way); <= 0 always builds the LBVH.
 *   "grid_prop_blocks": blocks per grid propagation launch, 8..65536 (rounded
 *   down to a multiple of 8; default 4096).
 *   "grid_batch_prop_blocks": the same for batched grid calls (several
 *   requests per launch; default 1536).
 *   "lbvh_visit_budget": LBVH node visits after which a nearest search stops
 *   in the per-query pass and finishes in a second pass that gives the query a
 *   whole wave, 16..2^31-1, or 0 (the default) for 15 x cbrt(triangles) per
 *   mesh.
 *   "grid_item_major_min": grid rounds after one that took at least this many
 *   cells propagate one lane per frontier cell rather than one lane per
 *   (cell, neighbour) pair, >= 0 (0: every round; default 2^20). */
int sdfx_set_tuning(const char* key, int64_t value);
int64_t sdfx_get_tuning(const char* key);

/* ---- Measurement (no counterpart in the reference; for benchmark and
 * profiling tools). While enabled, every kernel launch is bracketed by a HIP
 * event pair on its stream. sdfx_profile_enable(on) also clears the totals.
 * sdfx_profile_read fills up to max_entries rows (name: 64 bytes each, total
 * milliseconds, launch count) and returns the number of kernel names seen. */
void sdfx_profile_enable(int on);
int sdfx_profile_read(int max_entries, char* names, double* total_ms, uint64_t* launches);
/* Work counters accumulated while profiling (e.g. "grid.propagate_distance_evals"). */
uint64_t sdfx_profile_counter(const char* name);
/* Narrows the profiling started by sdfx_profile_enable(1): only kernels whose
 * name starts with kernel_prefix get an event pair ("" or NULL: all), and work
 * counters run only when counters != 0. Fewer event pairs disturb the launch
 * stream less, so one kernel's average duration is closest to a profiler's. */
void sdfx_profile_select(const char* kernel_prefix, int counters);

/* ---- Served form (no counterpart in the reference, which answers one call
 * at a time on the caller's thread; for callers that answer a stream of
 * requests against meshes kept on the device).
 *
 * A server keeps meshes resident (uploaded once: triangles in topology order
 * for grid requests, an LBVH for point requests) and runs queued requests on
 * `workers` threads, each with its own stream. A worker that comes free takes
 * every queued request that fits the free part of the HBM budget, first come
 * first served, and runs them as one batch: the grid requests of each sign
 * method in one batched grid run, the point requests of each query mode in one
 * multi-mesh LBVH launch. Requests submitted while batches run join the next
 * batch (continuous batching); batches of different workers overlap on the
 * device. Each answer is bit-identical to the request served alone.
 *
 * Budget: resident meshes count 256 bytes per triangle; a request in flight
 * counts its scratch (about 28 bytes per grid cell plus 40 per triangle of its
 * mesh, 24 bytes per query point). Query points and outputs are the caller's
 * device buffers and are not counted; they must stay valid until the request
 * has been waited for.
 *
 * Point requests answer query_sdf (lib.rs:463-485) on an acceleration of the
 * given method; like it, they make no finiteness check of the points. Grid
 * requests answer generate_grid_sdf (generate/grid.rs:265-273) on the mesh. */
typedef struct sdfx_server sdfx_server;

typedef struct sdfx_server_stats {
    uint64_t requests_done;
    uint64_t batches;             /* batches run (a batch is one worker's turn) */
    uint64_t max_batch_requests;  /* most requests in one batch */
    uint64_t resident_bytes;      /* budget held by resident meshes */
    uint64_t inflight_bytes;      /* budget held by running requests */
    uint64_t peak_inflight_bytes;
    uint64_t meshes;
    uint64_t queued;              /* requests waiting for a worker */
} sdfx_server_stats;

/* Starts a server on the current device with `workers` (1..8) worker threads. */
sdfx_status sdfx_server_create(uint64_t hbm_budget_bytes, uint32_t workers, sdfx_server** server_out);

/* Serves every queued request, then stops the workers and frees the meshes. */
void sdfx_server_destroy(sdfx_server* server);

/* Uploads a mesh (device pointers, copied: the caller may free them after the
 * call) and builds its LBVH. SDFX_ERR_OVER_BUDGET when it cannot fit beside the
 * resident meshes; otherwise waits until the requests in flight leave room.
 * Topology and index errors as generate_grid_sdf reports them. */
sdfx_status sdfx_server_add_mesh(sdfx_server* server, const float* d_vertices, uint64_t vertex_count,
                                 const uint32_t* d_indices, uint64_t index_count, sdfx_topology topology,
                                 uint32_t* mesh_id_out);

/* One mesh of sdfx_server_add_meshes: device pointers in, id and status out. */
typedef struct sdfx_server_mesh {
    const float* vertices;
    uint64_t vertex_count;
    const uint32_t* indices; /* NULL: no index buffer */
    uint64_t index_count;
    sdfx_topology topology;
    uint32_t id;        /* set on success */
    sdfx_status status; /* set by the call */
} sdfx_server_mesh;

/* sdfx_server_add_mesh for many meshes, admitted in order while they fit the
 * budget, with their LBVHs built in one batched device build. Returns SDFX_OK
 * when all were added, else the first failing mesh's status (each mesh's own
 * status and id are filled in; sdfx_last_error names the mesh). */
sdfx_status sdfx_server_add_meshes(sdfx_server* server, sdfx_server_mesh* meshes, uint64_t count);

/* Drops a mesh from the server (requests already queued for it still run). */
sdfx_status sdfx_server_remove_mesh(sdfx_server* server, uint32_t mesh_id);

/* Queues a point request: d_out[i] = the signed distance of d_query_points[i]
 * (device pointers). The request's status (sdfx_generate_sdf's) comes from
 * sdfx_server_wait. */
sdfx_status sdfx_server_submit_points(sdfx_server* server, uint32_t mesh_id, const float* d_query_points,
                                      uint64_t query_count, sdfx_acceleration_method method, sdfx_sign_method sign,
                                      float* d_out, uint64_t* ticket_out);

/* Queues a grid request on the mesh (d_out: device pointer, z fastest). */
sdfx_status sdfx_server_submit_grid(sdfx_server* server, uint32_t mesh_id, const float first_cell[3],
                                    const float cell_size[3], const uint64_t cell_count[3], sdfx_sign_method sign,
                                    float* d_out, uint64_t* ticket_out);

/* Blocks until the request is done; returns its status (its message through
 * sdfx_last_error) and, when latency_ms is not NULL, the milliseconds from
 * submission to completion. A ticket can be waited for once. */
sdfx_status sdfx_server_wait(sdfx_server* server, uint64_t ticket, double* latency_ms);

/* Cancels a request that is still queued: *cancelled_out = 1 and its wait
 * returns SDFX_ERR_CANCELLED. A request already running or done is left as it
 * is (*cancelled_out = 0). */
sdfx_status sdfx_server_cancel(sdfx_server* server, uint64_t ticket, int* cancelled_out);

/* *done_out = 1 when the request is done (it must still be waited for). */
sdfx_status sdfx_server_poll(sdfx_server* server, uint64_t ticket, int* done_out);

sdfx_status sdfx_server_get_stats(sdfx_server* server, sdfx_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* SDFX_H */
