/*
 * sdf_oracle.h — CPU restatement of mesh_to_sdf's hot path (generate_sdf and
 * generate_grid_sdf), written in plain C from the reference's Rust sources.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker for the HIP
 * product path and the CPU baseline timed beside it. Only tests/, the smoke()
 * entry of __graft_entry__.py and bench.py's cpu_baseline leg may load it.
 * The product (libsdfx) never links, loads or falls back to it.
 *
 * Numerics: FP32 evaluated in the reference's operation order, compiled with
 * -ffp-contract=off -fno-fast-math on x86-64 (SSE, no excess precision), so
 * every expression rounds exactly as the Rust code does (Rust never contracts
 * a*b+c into an FMA and its sqrt and / are IEEE correctly rounded).
 *
 * Status codes are numerically identical to include/sdfx.h's sdfx_status.
 */
#ifndef SDF_ORACLE_H
#define SDF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_EMPTY_RESULT = 1,
    ORC_ERR_NON_FINITE_QUERY = -1,
    ORC_ERR_INDEX_COUNT = -2,
    ORC_ERR_VERTEX_COUNT = -3,
    ORC_ERR_INDEX_OUT_OF_RANGE = -4,
    ORC_ERR_EMPTY_MESH = -5,
    ORC_ERR_NAN_DISTANCE = -6,
    ORC_ERR_INVALID_ARGUMENT = -7,
    ORC_ERR_INVALID_GRID = -10
};

enum { ORC_TOPO_LIST = 0, ORC_TOPO_STRIP = 1 };
enum { ORC_SIGN_RAYCAST = 0, ORC_SIGN_NORMAL = 1 };
enum { ORC_ACCEL_NONE = 0, ORC_ACCEL_BVH = 1, ORC_ACCEL_RTREE = 2, ORC_ACCEL_RTREE_BVH = 3 };
/* grid propagation schedules */
enum { ORC_GRID_SEQUENTIAL = 0, ORC_GRID_THREADED = 1, ORC_GRID_FRONTIER = 2, ORC_GRID_EXACT = 3 };

/* ---- geometry (geo.rs) ---- */
void orc_triangle_bounding_box(const float* a, const float* b, const float* c, float* lo, float* hi);
void orc_closest_point_triangle(const float* p, const float* a, const float* b, const float* c, float* out);
void orc_closest_point_segment(const float* p, const float* a, const float* b, float* out);
float orc_point_triangle_distance(const float* p, const float* a, const float* b, const float* c);
float orc_point_triangle_distance2(const float* p, const float* a, const float* b, const float* c);
float orc_point_triangle_signed_distance(const float* p, const float* a, const float* b, const float* c);
int orc_ray_triangle_intersection_aligned(const float* origin, const float* a, const float* b, const float* c,
                                          int axis, float* t_out);
int orc_ray_triangle_intersection_generic(const float* origin, const float* dir, const float* a, const float* b,
                                          const float* c, float* t_out);
float orc_baseline_point_triangle_distance(const float* p, const float* a, const float* b, const float* c);

/* ---- comparisons (lib.rs, float-cmp 0.9.0) ---- */
int orc_approx_eq_f32(float a, float b, int32_t ulps, float epsilon);
/* -1 Less, 0 Equal, 1 Greater, 2 = the reference panics ("NaN distance") */
int orc_compare_distances(float a, float b);

/* ---- topology (lib.rs:193-215) ---- */
int orc_decode_topology(uint64_t n_vertices, const uint32_t* indices, uint64_t n_indices, int topology,
                        uint32_t* tris_out, uint64_t tris_capacity, uint64_t* n_tris_out);

/* ---- grid arithmetic (grid.rs) ---- */
void orc_grid_from_bounding_box(const float* bmin, const float* bmax, const uint64_t* count, float* first,
                                float* size);
void orc_grid_cell_center(const float* first, const float* size, const uint64_t* cell, float* out);
/* returns 1 for SnapResult::Inside, 0 for Outside; -1 when the reference panics */
int orc_grid_snap(const float* first, const float* size, const uint64_t* count, const float* p, uint64_t* cell);

/* ---- entry points ---- */
int orc_generate_sdf(const float* vertices, uint64_t n_vertices, const uint32_t* indices, uint64_t n_indices,
                     int topology, const float* queries, uint64_t n_queries, int accel, int sign, float* out,
                     int n_threads, uint64_t* out_len);

typedef struct {
    uint64_t seeded_cells;   /* cells holding a value after seeding */
    uint64_t seed_updates;   /* successful preheap updates */
    uint64_t pops;           /* heap pops during propagation */
    uint64_t rays;           /* rays cast */
    uint64_t ray_candidates; /* candidates returned by BVH traversal */
    uint64_t rounds;         /* frontier schedule: propagation rounds */
    uint64_t evals;          /* frontier schedule: neighbour evaluations */
} orc_grid_stats;

/* schedule ORC_GRID_SEQUENTIAL: preheap in triangle order, the sorted heap split
 * round-robin into n_heaps heaps run one after another (one legal schedule of the
 * reference's concurrent threads; n_heaps=1 is the single-thread reference).
 * schedule ORC_GRID_THREADED: n_heaps OS threads run concurrently, as the
 * reference does (result depends on timing).
 * schedule ORC_GRID_FRONTIER: not one of the reference's schedules. The same
 * relaxation rule (a cell's triangle is tried on its 27 neighbours, kept when
 * compare_distances calls it Less than the stored value) run as rounds, each
 * cell holding the minimum (distance, triangle) under compare_distances' order
 * with the lower triangle first among equal distances (that order is a strict
 * total order, so the minimum does not depend on the order in which the
 * candidates arrive). A cell is flagged when its value improves (seeded
 * cells start flagged). Round r takes the flagged cells whose distance is at
 * most T_r = m_r + band, where m_r is the smallest distance of any flagged cell
 * and band = ORC_GRID_BAND_CELLS * max|cell_size| (both in FP32), snapshots
 * their (cell, triangle), clears their flags and tries each on its neighbours;
 * the other flagged cells wait. The neighbours are the reference's 26 plus,
 * along each axis k whose cells are at least 1.5 times thinner than the
 * largest (orc_grid_reach(max|cell_size|, |cell_size_k|) = 2), the two cells
 * two steps away on that axis: a front then moves about as far along a thin
 * axis per round as along the others (on the 128^3 bench grid, whose z cells
 * are 3.9 times thinner, 69 rounds become 52 and 4.09M items 3.33M). More
 * candidates only move values closer to the exact distance, and the end
 * state still satisfies the reference's fixpoint over the 26. Taking the nearest cells first, like the
 * reference's heap, keeps a cell from being improved many times by far
 * triangles before its near ones arrive. A round after one that took fewer
 * than max(ORC_GRID_RELEASE_MIN, cells / ORC_GRID_RELEASE_DIV) cells takes
 * every flagged cell instead (T_r = infinity): once the front is thin, waiting
 * only adds rounds. When such a thin round happens with m_r beyond
 * ORC_GRID_FAR_CELLS * max|cell_size| (FP32), every later round takes every
 * flagged cell: in the far field the front carries mostly final values, so the
 * band there only adds rounds. So does a second thin round wherever it
 * happens: once a released round has been followed by a banded round that
 * took almost nothing, the band only alternates such rounds with released
 * ones (on the 128^3 bench grid, 52 rounds become 45, for 0.1% more
 * evaluations and 96 fewer cells beyond 1e-5 x diagonal of the exact
 * distance). It is deterministic and is the schedule the
 * device pipeline runs, so it pins the device kernels bit-exactly; its
 * distance from the reference's own schedules is measured separately.
 * schedule ORC_GRID_EXACT: not one of the reference's schedules either: the
 * value the flood fill approaches, with no schedule. Raycast: each cell holds
 * the FP32 minimum over all triangles of geo.rs's distance from its centre
 * (the triangle of the smallest d^2, the lowest index among equal ones),
 * signed by the same grid raycasts (generate/grid.rs:568-684); Normal: the
 * compare_distances fold of geo.rs's signed distances over all triangles, the
 * value Bvh(Normal) gives the cell centre as a query point
 * (generic/bvh.rs:117-129). tri_out receives the nearest triangle (Raycast)
 * or UINT32_MAX (Normal). */
#define ORC_GRID_BAND_CELLS 3.0f
#define ORC_GRID_RELEASE_MIN 1024u
#define ORC_GRID_RELEASE_DIV 128u
#define ORC_GRID_FAR_CELLS 8.0f
/* 2 when max_size / abs_size >= 1.5 in FP32 (also false for NaN), else 1. */
unsigned orc_grid_reach(float max_size, float abs_size);
int orc_generate_grid_sdf(const float* vertices, uint64_t n_vertices, const uint32_t* indices, uint64_t n_indices,
                          int topology, const float* first_cell, const float* cell_size,
                          const uint64_t* cell_count, int sign, float* out, int n_heaps, int schedule,
                          float* seed_out, orc_grid_stats* stats);

/* orc_generate_grid_sdf plus, when tri_out is not NULL, the triangle each
 * cell's value is the distance to (the topology-order index of the triangle
 * the last improving state carried; UINT32_MAX for a cell with no value). */
int orc_generate_grid_sdf_ex(const float* vertices, uint64_t n_vertices, const uint32_t* indices, uint64_t n_indices,
                             int topology, const float* first_cell, const float* cell_size,
                             const uint64_t* cell_count, int sign, float* out, uint32_t* tri_out, int n_heaps,
                             int schedule, float* seed_out, orc_grid_stats* stats);

/* Counts the violations of the state every schedule of the reference's
 * propagation reaches (generate/grid.rs:495-558), given each cell's value and
 * triangle: counts[0] cells whose value is not their triangle's distance bit
 * for bit, counts[1] (cell, neighbour) pairs where the cell's triangle gives
 * the neighbour a value compare_distances calls Less than the neighbour's,
 * counts[2] cells without a triangle, counts[3] of those whose value is not
 * f32::MAX. For Raycast, |value| is compared (the sign comes from the rays). */
int orc_grid_fixpoint_check(const float* vertices, uint64_t n_vertices, const uint32_t* indices,
                            uint64_t n_indices, int topology, const float* first_cell, const float* cell_size,
                            const uint64_t* cell_count, int sign, const float* values, const uint32_t* tri_of,
                            int n_threads, uint64_t* counts);

/* Test-only fault injection into ORC_GRID_FRONTIER, so tests can show that
 * the fixpoint check rejects a wrong schedule: ORC_FAULT_SKIP_DIRECTION never
 * offers a cell's triangle to its +x neighbour, ORC_FAULT_DROP_DEFERRED drops
 * the flagged cells beyond a round's band instead of keeping them,
 * ORC_FAULT_PACKED_ORDER orders Normal-mode values by the packed key alone
 * (|d|, then negative after positive) instead of compare_distances, which
 * prefers a positive value over a negative one it approximately equals.
 * ORC_FAULT_NO_SECOND_RELEASE leaves out the release after a second thin
 * round (a schedule change, not a fault: tests use it to show which grids the
 * rule changes). */
enum { ORC_FAULT_NONE = 0, ORC_FAULT_SKIP_DIRECTION = 1, ORC_FAULT_DROP_DEFERRED = 2, ORC_FAULT_PACKED_ORDER = 3,
       ORC_FAULT_BOUND_MARGIN_SIGN = 4, ORC_FAULT_NO_SECOND_RELEASE = 5 };
void orc_set_fault(int kind);

/* Test-only check of the device propagation's supporting-plane skip
 * (plane_bound in sdfx_grid.hip), restated in FP32 with the same operations.
 * With mode 1 (direction words as improvements write them: the quantised
 * vector from the item cell's closest point on its triangle to its centre) or
 * mode 2 (a pseudo-random word per item), every ORC_GRID_FRONTIER propagation
 * pair whose neighbour holds a key from another triangle is tested: counts[0]
 * such pairs, counts[1] those the bound skips, counts[2] skipped pairs whose
 * distance would still have lowered the neighbour's key (must be 0; the
 * schedule itself evaluates every pair, so results are unchanged). Mode 0:
 * off. ORC_FAULT_BOUND_MARGIN_SIGN adds the bound's rounding margin instead of
 * subtracting it, which the check must catch. */
void orc_set_bound_check(int mode);
void orc_bound_check_counts(uint64_t* counts);

/* Vectors from each query to its closest point on the nearest triangle, plus the
 * 3-ray inside vote (generate/generic/rtree_bvh.rs:217-276). */
int orc_query_vector_to_closest_point(const float* vertices, uint64_t n_vertices, const uint32_t* indices,
                                      uint64_t n_indices, int topology, const float* queries, uint64_t n_queries,
                                      float* vec_out, uint8_t* inside_out, int n_threads);

/* ---- helpers for tests ---- */
/* exact FP32 nearest triangle (lowest index on ties) by brute force */
void orc_nearest_bruteforce(const float* vertices, const uint32_t* tris, uint64_t n_tris, const float* points,
                            uint64_t n_points, float* d2_out, uint32_t* tri_out, int n_threads);
/* the same through the oracle's BVH (identical results, for large point sets) */
void orc_nearest_bvh(const float* vertices, const uint32_t* tris, uint64_t n_tris, const float* points,
                     uint64_t n_points, float* d2_out, uint32_t* tri_out, int n_threads);
/* FP64 brute-force unsigned distance (Ericson's closest point evaluated in double) */
void orc_fp64_distance(const float* vertices, const uint32_t* tris, uint64_t n_tris, const float* points,
                       uint64_t n_points, double* out, int n_threads);
/* number of +axis ray crossings per point over all triangles (brute force) */
void orc_ray_parity_bruteforce(const float* vertices, const uint32_t* tris, uint64_t n_tris, const float* points,
                               uint64_t n_points, uint8_t* parity_out /* 3 per point */, int n_threads);

/* mesh_check.rs:24-81 (returns 0 when Ok, else an error kind 1..5) */
int orc_check_mesh_triangle_list(const float* vertices, uint64_t n_vertices, const uint32_t* faces,
                                 uint64_t n_faces_idx);
/* The same, with the failing vertex index (kind 2), face index (3) or edge
 * index 3 * face + k (4, 5) in *where_out. */
int orc_check_mesh_triangle_list_at(const float* vertices, uint64_t n_vertices, const uint32_t* faces,
                                    uint64_t n_faces_idx, uint64_t* where_out);

#ifdef __cplusplus
}
#endif
#endif
