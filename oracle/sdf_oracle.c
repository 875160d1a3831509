/*
 * sdf_oracle.c — CPU restatement of mesh_to_sdf's hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and CPU baseline for the HIP
 * product path (see sdf_oracle.h). Citations are to the reference crate,
 * relative to mesh_to_sdf/src/ unless stated otherwise.
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math -fopenmp -fPIC -shared
 * (see oracle/Makefile). Every float expression below keeps the reference's
 * evaluation order; do not "simplify" them.
 */
#include "sdf_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float x, y, z;
} v3;

static inline v3 mk(float x, float y, float z) {
    v3 r = {x, y, z};
    return r;
}
static inline v3 ld(const float* p) { return mk(p[0], p[1], p[2]); }
static inline void st(float* p, v3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
static inline float get(v3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

/* point.rs:82-88 */
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
/* point.rs:91-97 */
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
/* point.rs:100-102 — ((x*x' + y*y') + z*z') */
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* point.rs:105-111 */
static inline v3 vcross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* point.rs:114-121 */
static inline float vlength(v3 a) { return sqrtf(vdot(a, a)); }
static inline float vdist(v3 a, v3 b) { return vlength(vsub(a, b)); }
/* point.rs:124-127 */
static inline float vdist2(v3 a, v3 b) {
    v3 d = vsub(a, b);
    return vdot(d, d);
}
/* point.rs:131-133 */
static inline v3 vmul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
/* point.rs:136-142 */
static inline v3 comp_div(v3 a, v3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
/* PartialEq on [f32; 3]: component-wise == */
static inline int veq(v3 a, v3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

/* ------------------------------------------------------------------ geo.rs */

/* geo.rs:4-22 */
static void bbox(v3 a, v3 b, v3 c, v3* lo, v3* hi) {
    const float eps = 0.0001f;
    v3 mn = mk(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z)));
    v3 mx = mk(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z)));
    v3 e = mk(eps, eps, eps);
    *lo = vsub(mn, e);
    *hi = vadd(mx, e);
}

/* geo.rs:141-151; f32::clamp keeps NaN */
static v3 closest_segment(v3 p, v3 a, v3 b) {
    v3 ab = vsub(b, a);
    float m = vdot(ab, ab);
    v3 ap = vsub(p, a);
    float s = vdot(ab, ap) / m;
    if (s < 0.0f) s = 0.0f;
    else if (s > 1.0f) s = 1.0f;
    return vadd(a, vmul(ab, s));
}

/* geo.rs:70-138 (Embree/Ericson closest point, degenerate fallbacks first) */
static v3 closest_triangle(v3 p, v3 a, v3 b, v3 c) {
    int ab_eq = veq(a, b), bc_eq = veq(b, c), ac_eq = veq(a, c);
    if (ab_eq && bc_eq && ac_eq) return a;
    if (ab_eq) return closest_segment(p, a, c);
    if (bc_eq) return closest_segment(p, a, b);
    if (ac_eq) return closest_segment(p, a, b);

    v3 ab = vsub(b, a);
    v3 ac = vsub(c, a);
    v3 ap = vsub(p, a);
    float d1 = vdot(ab, ap);
    float d2 = vdot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;

    v3 bp = vsub(p, b);
    float d3 = vdot(ab, bp);
    float d4 = vdot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;

    v3 cp = vsub(p, c);
    float d5 = vdot(ab, cp);
    float d6 = vdot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;

    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float v = d1 / (d1 - d3);
        return vadd(a, vmul(ab, v));
    }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        float v = d2 / (d2 - d6);
        return vadd(a, vmul(ac, v));
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && d4 - d3 >= 0.0f && d5 - d6 >= 0.0f) {
        float v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        v3 bc = vsub(c, b);
        return vadd(b, vmul(bc, v));
    }
    float denom = 1.0f / (va + vb + vc);
    float v = vb * denom;
    float w = vc * denom;
    return vadd(vadd(a, vmul(ab, v)), vmul(ac, w));
}

/* geo.rs:26-30 */
static inline float tri_distance(v3 p, v3 a, v3 b, v3 c) { return vdist(p, closest_triangle(p, a, b, c)); }
/* geo.rs:33-37 */
static inline float tri_distance2(v3 p, v3 a, v3 b, v3 c) { return vdist2(p, closest_triangle(p, a, b, c)); }
/* geo.rs:43-56, normal from geo.rs:60-64; 0 maps to the negative branch */
static float tri_signed_distance(v3 p, v3 a, v3 b, v3 c) {
    v3 nearest = closest_triangle(p, a, b, c);
    v3 direction = vsub(p, nearest);
    v3 normal = vcross(vsub(b, a), vsub(c, a));
    float d = vdist(p, nearest);
    return vdot(direction, normal) > 0.0f ? d : -d;
}

/* geo.rs:165-216: the projection-plane axes rotate with the ray axis. */
static int ray_aligned(v3 o, v3 t0, v3 t1, v3 t2, int axis, float* t_out) {
    v3 e01 = vsub(t1, t0);
    v3 e12 = vsub(t2, t1);
    v3 e20 = vsub(t0, t2);
    v3 p0 = vsub(o, t0);
    v3 p1 = vsub(o, t1);
    v3 p2 = vsub(o, t2);
    /* geo.rs:181-195: (X: y,z,x) (Y: z,x,y) (Z: x,y,z) as (get_y, get_z, get_x) */
    int iy = axis == 0 ? 1 : (axis == 1 ? 2 : 0);
    int iz = axis == 0 ? 2 : (axis == 1 ? 0 : 1);
    int ix = axis;
    float w0 = get(p1, iz) * get(e12, iy) - get(p1, iy) * get(e12, iz);
    float w1 = get(p2, iz) * get(e20, iy) - get(p2, iy) * get(e20, iz);
    float w2 = get(p0, iz) * get(e01, iy) - get(p0, iy) * get(e01, iz);
    if ((w0 < 0.0f && w1 < 0.0f && w2 < 0.0f) || (w0 > 0.0f && w1 > 0.0f && w2 > 0.0f)) {
        float t = -(w0 * get(p0, ix) + w2 * get(p2, ix) + w1 * get(p1, ix)) / (w0 + w1 + w2);
        if (t > 0.0f) {
            if (t_out) *t_out = t;
            return 1;
        }
    }
    return 0;
}

void orc_triangle_bounding_box(const float* a, const float* b, const float* c, float* lo, float* hi) {
    v3 l, h;
    bbox(ld(a), ld(b), ld(c), &l, &h);
    st(lo, l);
    st(hi, h);
}
void orc_closest_point_triangle(const float* p, const float* a, const float* b, const float* c, float* out) {
    st(out, closest_triangle(ld(p), ld(a), ld(b), ld(c)));
}
void orc_closest_point_segment(const float* p, const float* a, const float* b, float* out) {
    st(out, closest_segment(ld(p), ld(a), ld(b)));
}
float orc_point_triangle_distance(const float* p, const float* a, const float* b, const float* c) {
    return tri_distance(ld(p), ld(a), ld(b), ld(c));
}
float orc_point_triangle_distance2(const float* p, const float* a, const float* b, const float* c) {
    return tri_distance2(ld(p), ld(a), ld(b), ld(c));
}
float orc_point_triangle_signed_distance(const float* p, const float* a, const float* b, const float* c) {
    return tri_signed_distance(ld(p), ld(a), ld(b), ld(c));
}
int orc_ray_triangle_intersection_aligned(const float* origin, const float* a, const float* b, const float* c,
                                          int axis, float* t_out) {
    return ray_aligned(ld(origin), ld(a), ld(b), ld(c), axis, t_out);
}

/* geo.rs:397-455 (the reference's test-only generic ray test) */
int orc_ray_triangle_intersection_generic(const float* origin, const float* dir, const float* a, const float* b,
                                          const float* c, float* t_out) {
    v3 o = ld(origin), d = ld(dir), v0 = ld(a), v1 = ld(b), v2 = ld(c);
    v3 n = vcross(vsub(v1, v0), vsub(v2, v0));
    float ndd = vdot(n, d);
    if (fabsf(ndd) < 0.00001f) return 0;
    float dd = -vdot(n, v0);
    float t = -(vdot(n, o) + dd) / ndd;
    if (t < 0.0f) return 0;
    v3 p = vadd(o, vmul(d, t));
    if (vdot(n, vcross(vsub(v1, v0), vsub(p, v0))) < 0.0f) return 0;
    if (vdot(n, vcross(vsub(v2, v1), vsub(p, v1))) < 0.0f) return 0;
    if (vdot(n, vcross(vsub(v0, v2), vsub(p, v2))) < 0.0f) return 0;
    if (t_out) *t_out = t;
    return 1;
}

/* geo.rs:383-393 */
static float point_segment_distance(v3 x0, v3 x1, v3 x2) {
    v3 dx = vsub(x2, x1);
    float m2 = vdot(dx, dx);
    float s12 = vdot(dx, vsub(x2, x0)) / m2;
    if (s12 < 0.0f) s12 = 0.0f;
    else if (s12 > 1.0f) s12 = 1.0f;
    return vdist(x0, vadd(vmul(x1, s12), vmul(x2, 1.0f - s12)));
}

/* geo.rs:329-380 (SDFGen-style baseline used by the reference's property test) */
float orc_baseline_point_triangle_distance(const float* p, const float* a, const float* b, const float* c) {
    v3 x0 = ld(p), x1 = ld(a), x2 = ld(b), x3 = ld(c);
    v3 x03 = vsub(x0, x3), x13 = vsub(x1, x3), x23 = vsub(x2, x3);
    float m13 = vdot(x13, x13), m23 = vdot(x23, x23), d = vdot(x13, x23);
    float invdet = 1.0f / fmaxf(m13 * m23 - d * d, 1e-30f);
    float aa = vdot(x13, x03), bb = vdot(x23, x03);
    float w23 = invdet * (m23 * aa - d * bb);
    float w31 = invdet * (m13 * bb - d * aa);
    float w12 = 1.0f - w23 - w31;
    if (w23 >= 0.0f && w31 >= 0.0f && w12 >= 0.0f) {
        v3 proj = vadd(vadd(vmul(x1, w23), vmul(x2, w31)), vmul(x3, w12));
        return vdist(x0, proj);
    }
    if (w23 > 0.0f) return fminf(point_segment_distance(x0, x1, x2), point_segment_distance(x0, x1, x3));
    if (w31 > 0.0f) return fminf(point_segment_distance(x0, x1, x2), point_segment_distance(x0, x2, x3));
    return fminf(point_segment_distance(x0, x1, x3), point_segment_distance(x0, x2, x3));
}

/* ------------------------------------------------------ compare_distances */

/* float-cmp 0.9.0 ApproxEq for f32: exact, then |a-b| <= eps, then
 * saturating |ulps(a,b)| <= ulps where ulps is the wrapping i32 difference of
 * the bit patterns. */
int orc_approx_eq_f32(float a, float b, int32_t ulps, float epsilon) {
    if (a == b) return 1;
    if (fabsf(a - b) <= epsilon) return 1;
    int32_t ai, bi;
    memcpy(&ai, &a, 4);
    memcpy(&bi, &b, 4);
    int32_t diff = (int32_t)((uint32_t)ai - (uint32_t)bi);
    int32_t ad = diff == INT32_MIN ? INT32_MAX : (diff < 0 ? -diff : diff);
    return ad <= ulps;
}

static inline int sign_negative(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    return (int)(u >> 31);
}

/* lib.rs:315-332 */
int orc_compare_distances(float a, float b) {
    float aa = fabsf(a), bb = fabsf(b);
    if (orc_approx_eq_f32(aa, bb, 2, 1e-6f)) {
        int an = sign_negative(a), bn = sign_negative(b);
        if (an && !bn) return 1;
        if (!an && bn) return -1;
    }
    if (aa != aa || bb != bb) return 2;
    return aa < bb ? -1 : (aa > bb ? 1 : 0);
}

/* ----------------------------------------------------------------- topology */

/* lib.rs:193-215: TriangleList chunks of 3 (asserting len % 3 == 0),
 * TriangleStrip sliding windows with no winding alternation; an index beyond
 * the vertex array panics when the triangle is read. */
int orc_decode_topology(uint64_t nv, const uint32_t* idx, uint64_t ni, int topo, uint32_t* tris,
                        uint64_t cap, uint64_t* n_tris) {
    uint64_t nt = 0;
    if (topo == ORC_TOPO_LIST) {
        if (idx) {
            if (ni % 3 != 0) return ORC_ERR_INDEX_COUNT;
            nt = ni / 3;
        } else {
            if (nv % 3 != 0) return ORC_ERR_VERTEX_COUNT;
            nt = nv / 3;
        }
    } else if (topo == ORC_TOPO_STRIP) {
        uint64_t n = idx ? ni : nv;
        nt = n >= 3 ? n - 2 : 0;
    } else {
        return ORC_ERR_INVALID_ARGUMENT;
    }
    if (n_tris) *n_tris = nt;
    if (!tris) return ORC_OK;
    if (cap < nt) return ORC_ERR_INVALID_ARGUMENT;
    for (uint64_t t = 0; t < nt; ++t) {
        uint64_t base = topo == ORC_TOPO_LIST ? 3 * t : t;
        for (int k = 0; k < 3; ++k) {
            uint64_t v = idx ? (uint64_t)idx[base + k] : base + k;
            if (v >= nv) return ORC_ERR_INDEX_OUT_OF_RANGE;
            tris[3 * t + k] = (uint32_t)v;
        }
    }
    return ORC_OK;
}

/* --------------------------------------------------------------- BVH (CPU) */
/* A binary BVH with one triangle per leaf over the padded boxes of geo.rs:4-22,
 * built by binned SAH like the bvh 0.10.0 crate's builder. Only its
 * conservativeness matters for results: ray traversal returns every triangle
 * whose padded box the ray meets, nearest search returns the exact minimum. */
typedef struct {
    float lo[3], hi[3];
    int32_t left, right; /* leaf: left = -1 - triangle, right = -1 */
} onode;

typedef struct {
    onode* nodes;
    uint64_t n_nodes;
} obvh;

typedef struct {
    float lo[3], hi[3], c[3];
} obox;

static void box_union(float* lo, float* hi, const float* l2, const float* h2) {
    for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(lo[k], l2[k]);
        hi[k] = fmaxf(hi[k], h2[k]);
    }
}
static double box_area(const float* lo, const float* hi) {
    double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

static int32_t bvh_build_rec(obvh* b, const obox* boxes, uint32_t* ids, uint64_t n) {
    int32_t me = (int32_t)b->n_nodes++;
    onode* nd = &b->nodes[me];
    for (int k = 0; k < 3; ++k) {
        nd->lo[k] = FLT_MAX;
        nd->hi[k] = -FLT_MAX;
    }
    float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint64_t i = 0; i < n; ++i) {
        box_union(nd->lo, nd->hi, boxes[ids[i]].lo, boxes[ids[i]].hi);
        box_union(clo, chi, boxes[ids[i]].c, boxes[ids[i]].c);
    }
    if (n == 1) {
        nd->left = -1 - (int32_t)ids[0];
        nd->right = -1;
        return me;
    }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
        if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
    uint64_t mid = n / 2;
    float ext = chi[axis] - clo[axis];
    if (ext > 0.0f) {
        enum { NB = 16 };
        uint64_t cnt[NB] = {0};
        float blo[NB][3], bhi[NB][3];
        for (int i = 0; i < NB; ++i)
            for (int k = 0; k < 3; ++k) {
                blo[i][k] = FLT_MAX;
                bhi[i][k] = -FLT_MAX;
            }
        for (uint64_t i = 0; i < n; ++i) {
            int bi = (int)((boxes[ids[i]].c[axis] - clo[axis]) / ext * NB);
            if (bi >= NB) bi = NB - 1;
            if (bi < 0) bi = 0;
            cnt[bi]++;
            box_union(blo[bi], bhi[bi], boxes[ids[i]].lo, boxes[ids[i]].hi);
        }
        double best = 1e300;
        int best_split = -1;
        for (int s = 1; s < NB; ++s) {
            float llo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, lhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            float rlo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, rhi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            uint64_t nl = 0, nr = 0;
            for (int i = 0; i < s; ++i)
                if (cnt[i]) {
                    nl += cnt[i];
                    box_union(llo, lhi, blo[i], bhi[i]);
                }
            for (int i = s; i < NB; ++i)
                if (cnt[i]) {
                    nr += cnt[i];
                    box_union(rlo, rhi, blo[i], bhi[i]);
                }
            if (!nl || !nr) continue;
            double cost = nl * box_area(llo, lhi) + nr * box_area(rlo, rhi);
            if (cost < best) {
                best = cost;
                best_split = s;
            }
        }
        if (best_split > 0) {
            uint64_t lo_i = 0, hi_i = n;
            while (lo_i < hi_i) {
                int bi = (int)((boxes[ids[lo_i]].c[axis] - clo[axis]) / ext * NB);
                if (bi >= NB) bi = NB - 1;
                if (bi < 0) bi = 0;
                if (bi < best_split) {
                    lo_i++;
                } else {
                    uint32_t tmp = ids[lo_i];
                    ids[lo_i] = ids[--hi_i];
                    ids[hi_i] = tmp;
                }
            }
            if (lo_i > 0 && lo_i < n) mid = lo_i;
        }
    }
    int32_t l = bvh_build_rec(b, boxes, ids, mid);
    int32_t r = bvh_build_rec(b, boxes, ids + mid, n - mid);
    b->nodes[me].left = l;
    b->nodes[me].right = r;
    return me;
}

static int bvh_build(obvh* b, const float* verts, const uint32_t* tris, uint64_t nt) {
    b->nodes = NULL;
    b->n_nodes = 0;
    if (nt == 0) return 0;
    obox* boxes = (obox*)malloc(sizeof(obox) * nt);
    uint32_t* ids = (uint32_t*)malloc(sizeof(uint32_t) * nt);
    b->nodes = (onode*)malloc(sizeof(onode) * (2 * nt));
    if (!boxes || !ids || !b->nodes) {
        free(boxes);
        free(ids);
        free(b->nodes);
        b->nodes = NULL;
        return -1;
    }
    for (uint64_t t = 0; t < nt; ++t) {
        v3 lo, hi;
        bbox(ld(verts + 3 * tris[3 * t]), ld(verts + 3 * tris[3 * t + 1]), ld(verts + 3 * tris[3 * t + 2]), &lo,
             &hi);
        st(boxes[t].lo, lo);
        st(boxes[t].hi, hi);
        for (int k = 0; k < 3; ++k) boxes[t].c[k] = 0.5f * (boxes[t].lo[k] + boxes[t].hi[k]);
        ids[t] = (uint32_t)t;
    }
    bvh_build_rec(b, boxes, ids, nt);
    free(boxes);
    free(ids);
    return 0;
}

static void bvh_free(obvh* b) {
    free(b->nodes);
    b->nodes = NULL;
}

/* Count aligned-ray hits (geo.rs:165-216) over the triangles whose padded box the
 * +axis ray from o meets: perpendicular coordinates inside, box not behind. */
static int bvh_ray_hits(const obvh* b, const float* verts, const uint32_t* tris, v3 o, int axis,
                        float* ts, int cap, uint64_t* candidates) {
    if (!b->n_nodes) return 0;
    int32_t stack[128];
    int sp = 0, nh = 0;
    stack[sp++] = 0;
    float oc[3] = {o.x, o.y, o.z};
    int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    while (sp) {
        const onode* nd = &b->nodes[stack[--sp]];
        if (!(nd->lo[a1] <= oc[a1] && oc[a1] <= nd->hi[a1] && nd->lo[a2] <= oc[a2] && oc[a2] <= nd->hi[a2] &&
              nd->hi[axis] >= oc[axis]))
            continue;
        if (nd->right < 0) {
            uint32_t t = (uint32_t)(-1 - nd->left);
            if (candidates) (*candidates)++;
            float tt;
            if (ray_aligned(o, ld(verts + 3 * tris[3 * t]), ld(verts + 3 * tris[3 * t + 1]),
                            ld(verts + 3 * tris[3 * t + 2]), axis, &tt)) {
                if (ts && nh < cap) ts[nh] = tt;
                nh++;
            }
        } else {
            stack[sp++] = nd->left;
            stack[sp++] = nd->right;
        }
    }
    return nh;
}

static inline float box_d2(const onode* nd, v3 p) {
    float dx = fmaxf(fmaxf(nd->lo[0] - p.x, p.x - nd->hi[0]), 0.0f);
    float dy = fmaxf(fmaxf(nd->lo[1] - p.y, p.y - nd->hi[1]), 0.0f);
    float dz = fmaxf(fmaxf(nd->lo[2] - p.z, p.z - nd->hi[2]), 0.0f);
    return dx * dx + dy * dy + dz * dz;
}

/* Exact nearest triangle by FP32 d^2 (geo.rs:33-37), lowest triangle index on
 * exact ties. Boxes are only pruned when clearly farther (a 2^-16 relative
 * margin covers their rounding), so the result equals brute force. */
static void bvh_nearest(const obvh* b, const float* verts, const uint32_t* tris, v3 p, float* best_d2,
                        uint32_t* best_t) {
    float bd = INFINITY;
    uint32_t bt = UINT32_MAX;
    int32_t stack[128];
    int sp = 0;
    if (b->n_nodes) stack[sp++] = 0;
    while (sp) {
        const onode* nd = &b->nodes[stack[--sp]];
        if (box_d2(nd, p) * (1.0f - 1.0f / 65536.0f) > bd) continue;
        if (nd->right < 0) {
            uint32_t t = (uint32_t)(-1 - nd->left);
            float d2 = tri_distance2(p, ld(verts + 3 * tris[3 * t]), ld(verts + 3 * tris[3 * t + 1]),
                                     ld(verts + 3 * tris[3 * t + 2]));
            if (d2 < bd || (d2 == bd && t < bt)) {
                bd = d2;
                bt = t;
            }
        } else {
            const onode* l = &b->nodes[nd->left];
            const onode* r = &b->nodes[nd->right];
            float dl = box_d2(l, p), dr = box_d2(r, p);
            if (dl <= dr) {
                stack[sp++] = nd->right;
                stack[sp++] = nd->left;
            } else {
                stack[sp++] = nd->left;
                stack[sp++] = nd->right;
            }
        }
    }
    *best_d2 = bd;
    *best_t = bt;
}

/* 3-ray vote of generic/bvh.rs:141-176 and generic/rtree_bvh.rs:163-198:
 * inside iff at least two of the +X/+Y/+Z rays cross the mesh an odd number of times. */
static int vote_inside(const obvh* b, const float* verts, const uint32_t* tris, v3 p) {
    int insides = 0;
    for (int axis = 0; axis < 3; ++axis)
        if (bvh_ray_hits(b, verts, tris, p, axis, NULL, 0, NULL) % 2 == 1) insides++;
    return insides > 1;
}

/* ------------------------------------------------------------ entry points */

static int queries_finite(const float* q, uint64_t nq) {
    for (uint64_t i = 0; i < 3 * nq; ++i)
        if (!isfinite(q[i])) return 0;
    return 1;
}

static int decode_alloc(uint64_t nv, const uint32_t* idx, uint64_t ni, int topo, uint32_t** tris, uint64_t* nt) {
    int s = orc_decode_topology(nv, idx, ni, topo, NULL, 0, nt);
    if (s) return s;
    *tris = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (*nt + 1));
    if (!*tris) return ORC_ERR_INVALID_ARGUMENT;
    s = orc_decode_topology(nv, idx, ni, topo, *tris, *nt, nt);
    if (s) {
        free(*tris);
        *tris = NULL;
    }
    return s;
}

#define TRI_A(t) ld(verts + 3 * tris[3 * (t)])
#define TRI_B(t) ld(verts + 3 * tris[3 * (t) + 1])
#define TRI_C(t) ld(verts + 3 * tris[3 * (t) + 2])

/* lib.rs:364-391 dispatching to generic/{default,bvh,rtree,rtree_bvh}.rs */
int orc_generate_sdf(const float* verts, uint64_t nv, const uint32_t* idx, uint64_t ni, int topo,
                     const float* queries, uint64_t nq, int accel, int sign, float* out, int n_threads,
                     uint64_t* out_len) {
    if (out_len) *out_len = 0;
    if (n_threads <= 0) n_threads = 1;
    /* lib.rs:374-376 */
    if (!queries_finite(queries, nq)) return ORC_ERR_NON_FINITE_QUERY;
    if (accel < ORC_ACCEL_NONE || accel > ORC_ACCEL_RTREE_BVH) return ORC_ERR_INVALID_ARGUMENT;
    if (sign != ORC_SIGN_RAYCAST && sign != ORC_SIGN_NORMAL) return ORC_ERR_INVALID_ARGUMENT;
    /* default.rs:38-41 decodes the topology per query: nothing is checked without queries */
    if (accel == ORC_ACCEL_NONE && nq == 0) return ORC_OK;

    uint32_t* tris = NULL;
    uint64_t nt = 0;
    int s = decode_alloc(nv, idx, ni, topo, &tris, &nt);
    if (s) return s;

    int status = ORC_OK;
    if (accel == ORC_ACCEL_NONE) {
        /* default.rs:22-85: a fold over all triangles in order per query. */
        int nan_hit = 0;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 64) reduction(| : nan_hit)
        for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
            v3 p = ld(queries + 3 * qi);
            float mind = FLT_MAX;
            uint32_t count = 0;
            for (uint64_t t = 0; t < nt; ++t) {
                v3 a = TRI_A(t), b = TRI_B(t), c = TRI_C(t);
                if (sign == ORC_SIGN_RAYCAST) {
                    float d = tri_distance(p, a, b, c);
                    count += (uint32_t)ray_aligned(p, a, b, c, 0, NULL);
                    mind = fminf(mind, d);
                } else {
                    float d = tri_signed_distance(p, a, b, c);
                    int cmp = orc_compare_distances(d, mind);
                    if (cmp == 2) nan_hit = 1;
                    if (cmp == -1) mind = d;
                }
            }
            out[qi] = (count % 2 == 0) ? mind : -mind;
        }
        if (nan_hit) status = ORC_ERR_NAN_DISTANCE;
    } else {
        obvh bvh;
        if (bvh_build(&bvh, verts, tris, nt)) {
            free(tris);
            return ORC_ERR_INVALID_ARGUMENT;
        }
        if (accel == ORC_ACCEL_RTREE_BVH && nt == 0) {
            /* generic/rtree_bvh.rs:96-98, 124-126: an empty Vec */
            status = ORC_EMPTY_RESULT;
        } else if (nt == 0 && nq > 0) {
            /* Rtree: nearest_neighbor().unwrap() panics (generic/rtree.rs:148);
             * Bvh: the bvh crate has no root node to traverse. */
            status = ORC_ERR_EMPTY_MESH;
        } else {
            int nan_hit = 0;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 64) reduction(| : nan_hit)
            for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
                v3 p = ld(queries + 3 * qi);
                if (accel == ORC_ACCEL_BVH && sign == ORC_SIGN_NORMAL) {
                    /* generic/bvh.rs:117-129: compare_distances fold; the bvh
                     * crate's candidate order is replaced by triangle order. */
                    float mind = FLT_MAX;
                    for (uint64_t t = 0; t < nt; ++t) {
                        float d = tri_signed_distance(p, TRI_A(t), TRI_B(t), TRI_C(t));
                        int cmp = orc_compare_distances(mind, d);
                        if (cmp == 2) nan_hit = 1;
                        if (cmp == 1) mind = d;
                    }
                    out[qi] = mind;
                } else if (accel == ORC_ACCEL_BVH) {
                    /* generic/bvh.rs:131-176: min distance, then the 3-ray vote */
                    float d2;
                    uint32_t bt;
                    bvh_nearest(&bvh, verts, tris, p, &d2, &bt);
                    float d = tri_distance(p, TRI_A(bt), TRI_B(bt), TRI_C(bt));
                    out[qi] = vote_inside(&bvh, verts, tris, p) ? -d : d;
                } else if (accel == ORC_ACCEL_RTREE) {
                    /* generic/rtree.rs:147-154: nearest by d^2, then signed distance */
                    float d2;
                    uint32_t bt;
                    bvh_nearest(&bvh, verts, tris, p, &d2, &bt);
                    out[qi] = tri_signed_distance(p, TRI_A(bt), TRI_B(bt), TRI_C(bt));
                } else {
                    /* generic/rtree_bvh.rs:153-198 */
                    float d2;
                    uint32_t bt;
                    bvh_nearest(&bvh, verts, tris, p, &d2, &bt);
                    float d = tri_distance(p, TRI_A(bt), TRI_B(bt), TRI_C(bt));
                    out[qi] = vote_inside(&bvh, verts, tris, p) ? -d : d;
                }
            }
            if (nan_hit) status = ORC_ERR_NAN_DISTANCE;
        }
        bvh_free(&bvh);
    }
    free(tris);
    if (status == ORC_OK && out_len) *out_len = nq;
    return status;
}

/* generate/generic/rtree_bvh.rs:217-276 */
int orc_query_vector_to_closest_point(const float* verts, uint64_t nv, const uint32_t* idx, uint64_t ni, int topo,
                                      const float* queries, uint64_t nq, float* vec_out, uint8_t* inside_out,
                                      int n_threads) {
    if (n_threads <= 0) n_threads = 1;
    uint32_t* tris = NULL;
    uint64_t nt = 0;
    int s = decode_alloc(nv, idx, ni, topo, &tris, &nt);
    if (s) return s;
    if (nt == 0) {
        free(tris);
        return ORC_EMPTY_RESULT;
    }
    obvh bvh;
    bvh_build(&bvh, verts, tris, nt);
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 64)
    for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
        v3 p = ld(queries + 3 * qi);
        float d2;
        uint32_t bt;
        bvh_nearest(&bvh, verts, tris, p, &d2, &bt);
        v3 cp = closest_triangle(p, TRI_A(bt), TRI_B(bt), TRI_C(bt));
        st(vec_out + 3 * qi, vsub(cp, p));
        inside_out[qi] = (uint8_t)vote_inside(&bvh, verts, tris, p);
    }
    bvh_free(&bvh);
    free(tris);
    return ORC_OK;
}

/* --------------------------------------------------------------------- grid */

/* grid.rs:213-228 */
void orc_grid_from_bounding_box(const float* bmin, const float* bmax, const uint64_t* count, float* first,
                                float* size) {
    v3 fc = mk((float)count[0], (float)count[1], (float)count[2]);
    v3 cs = comp_div(vsub(ld(bmax), ld(bmin)), fc);
    st(size, cs);
    st(first, vadd(ld(bmin), vmul(cs, 0.5f)));
}

typedef struct {
    v3 first, size;
    uint64_t n[3];
} ogrid;

/* grid.rs:289-295 */
static inline v3 cell_center(const ogrid* g, uint64_t x, uint64_t y, uint64_t z) {
    return mk(g->first.x + (float)x * g->size.x, g->first.y + (float)y * g->size.y,
              g->first.z + (float)z * g->size.z);
}
/* grid.rs:276-278 */
static inline uint64_t cell_idx(const ogrid* g, uint64_t x, uint64_t y, uint64_t z) {
    return z + y * g->n[2] + x * g->n[1] * g->n[2];
}

void orc_grid_cell_center(const float* first, const float* size, const uint64_t* cell, float* out) {
    ogrid g = {ld(first), ld(size), {0, 0, 0}};
    st(out, cell_center(&g, cell[0], cell[1], cell[2]));
}

/* Rust `f32 as isize`: NaN -> 0, saturating at the isize range. */
static inline int64_t f2isize(float x) {
    if (x != x) return 0;
    if (x >= 9223372036854775808.0f) return INT64_MAX;
    if (x < -9223372036854775808.0f) return INT64_MIN;
    return (int64_t)x;
}
/* Rust `f32 as usize`: NaN and negatives -> 0, saturating above. */
static inline uint64_t f2usize(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)x;
}

/* grid.rs:264-273 then grid.rs:299-324; -1 when clamp(0, n-1) panics (n == 0) */
static int snap(const ogrid* g, v3 p, uint64_t* cell) {
    v3 bmin = vsub(g->first, vmul(g->size, 0.5f));
    v3 c = comp_div(vsub(p, bmin), g->size);
    int inside = 1;
    for (int k = 0; k < 3; ++k) {
        if (g->n[k] == 0) return -1;
        int64_t ci = f2isize(floorf(get(c, k)));
        int64_t hi = (int64_t)g->n[k] - 1;
        int64_t r = ci < 0 ? 0 : (ci > hi ? hi : ci);
        if (r != ci) inside = 0;
        cell[k] = (uint64_t)r;
    }
    return inside;
}

int orc_grid_snap(const float* first, const float* size, const uint64_t* count, const float* p, uint64_t* cell) {
    ogrid g = {ld(first), ld(size), {count[0], count[1], count[2]}};
    return snap(&g, ld(p), cell);
}

/* grid.rs:16-34: max-heap order; the state with the smaller distance (by
 * compare_distances), then the greater cell, then the greater triangle, wins. */
typedef struct {
    float d;
    uint32_t cell[3];
    uint32_t tri;
} ostate;

typedef struct {
    const uint32_t* tris;
} state_ctx;

static int state_cmp(const ostate* s, const ostate* o, const uint32_t* tris) {
    int c = orc_compare_distances(o->d, s->d);
    if (c == 2) c = 0;
    if (c) return c;
    for (int k = 0; k < 3; ++k)
        if (s->cell[k] != o->cell[k]) return s->cell[k] < o->cell[k] ? -1 : 1;
    for (int k = 0; k < 3; ++k) {
        uint32_t a = tris[3 * s->tri + k], b = tris[3 * o->tri + k];
        if (a != b) return a < b ? -1 : 1;
    }
    return 0;
}

/* std::collections::BinaryHeap's sift_up / sift_down_range /
 * sift_down_to_bottom, so the pop sequence matches even for the
 * non-transitive Normal-mode order. */
typedef struct {
    ostate* data;
    uint64_t len, cap;
    const uint32_t* tris;
} oheap;

static inline int le(const oheap* h, const ostate* a, const ostate* b) { return state_cmp(a, b, h->tris) != 1; }
static inline int ge(const oheap* h, const ostate* a, const ostate* b) { return state_cmp(a, b, h->tris) != -1; }
static inline int lt(const oheap* h, const ostate* a, const ostate* b) { return state_cmp(a, b, h->tris) == -1; }

static uint64_t sift_up(oheap* h, uint64_t start, uint64_t pos) {
    ostate elt = h->data[pos];
    while (pos > start) {
        uint64_t parent = (pos - 1) / 2;
        if (le(h, &elt, &h->data[parent])) break;
        h->data[pos] = h->data[parent];
        pos = parent;
    }
    h->data[pos] = elt;
    return pos;
}
static void sift_down_range(oheap* h, uint64_t pos, uint64_t end) {
    ostate elt = h->data[pos];
    uint64_t child = 2 * pos + 1;
    while (child <= (end >= 2 ? end - 2 : 0) && end >= 2) {
        child += le(h, &h->data[child], &h->data[child + 1]) ? 1 : 0;
        if (ge(h, &elt, &h->data[child])) {
            h->data[pos] = elt;
            return;
        }
        h->data[pos] = h->data[child];
        pos = child;
        child = 2 * pos + 1;
    }
    if (end >= 1 && child == end - 1 && lt(h, &elt, &h->data[child])) {
        h->data[pos] = h->data[child];
        pos = child;
    }
    h->data[pos] = elt;
}
static void sift_down_to_bottom(oheap* h, uint64_t pos) {
    uint64_t end = h->len, start = pos;
    ostate elt = h->data[pos];
    uint64_t child = 2 * pos + 1;
    while (end >= 2 && child <= end - 2) {
        child += le(h, &h->data[child], &h->data[child + 1]) ? 1 : 0;
        h->data[pos] = h->data[child];
        pos = child;
        child = 2 * pos + 1;
    }
    if (end >= 1 && child == end - 1) {
        h->data[pos] = h->data[child];
        pos = child;
    }
    h->data[pos] = elt;
    sift_up(h, start, pos);
}
static void heap_rebuild(oheap* h) {
    uint64_t n = h->len / 2;
    while (n > 0) {
        n--;
        sift_down_range(h, n, h->len);
    }
}
static int heap_push(oheap* h, ostate s) {
    if (h->len == h->cap) {
        uint64_t nc = h->cap ? 2 * h->cap : 1024;
        ostate* nd = (ostate*)realloc(h->data, nc * sizeof(ostate));
        if (!nd) return -1;
        h->data = nd;
        h->cap = nc;
    }
    h->data[h->len++] = s;
    sift_up(h, 0, h->len - 1);
    return 0;
}
static int heap_pop(oheap* h, ostate* out) {
    if (!h->len) return 0;
    ostate item = h->data[--h->len];
    if (h->len) {
        ostate top = h->data[0];
        h->data[0] = item;
        item = top;
        sift_down_to_bottom(h, 0);
    }
    *out = item;
    return 1;
}

static const uint32_t* g_sort_tris;
static int qsort_state_cmp(const void* a, const void* b) {
    return state_cmp((const ostate*)a, (const ostate*)b, g_sort_tris);
}

/* per-cell spin locks for the threaded schedule (parking_lot::RwLock in the reference) */
static inline void cell_lock(uint8_t* locks, uint64_t i) {
    while (__atomic_test_and_set(&locks[i], __ATOMIC_ACQUIRE)) {
    }
}
static inline void cell_unlock(uint8_t* locks, uint64_t i) { __atomic_clear(&locks[i], __ATOMIC_RELEASE); }

static inline float cell_distance(const ogrid* g, const float* verts, const uint32_t* tris, uint32_t t,
                                  const uint64_t* cell, int sign) {
    v3 p = cell_center(g, cell[0], cell[1], cell[2]);
    return sign == ORC_SIGN_RAYCAST ? tri_distance(p, TRI_A(t), TRI_B(t), TRI_C(t))
                                    : tri_signed_distance(p, TRI_A(t), TRI_B(t), TRI_C(t));
}

/* generate/grid.rs:383-457 for one triangle */
static uint64_t seed_triangle(const ogrid* g, const float* verts, const uint32_t* tris, uint32_t t, int sign,
                              float* pre_d, uint32_t* pre_t, uint8_t* locks) {
    v3 a = TRI_A(t), b = TRI_B(t), c = TRI_C(t);
    v3 lo, hi;
    bbox(a, b, c, &lo, &hi);
    uint64_t mn[3], mx[3];
    snap(g, lo, mn);
    snap(g, hi, mx);
    v3 mnf = cell_center(g, mn[0], mn[1], mn[2]);
    for (int i = 0; i < 3; ++i)
        if (mn[i] > 0 && get(mnf, i) > get(lo, i)) mn[i] -= 1;
    v3 mxf = cell_center(g, mx[0], mx[1], mx[2]);
    for (int i = 0; i < 3; ++i)
        if (mx[i] < g->n[i] - 1 && get(mxf, i) < get(hi, i)) mx[i] += 1;
    uint64_t updates = 0;
    for (uint64_t x = mn[0]; x <= mx[0]; ++x)
        for (uint64_t y = mn[1]; y <= mx[1]; ++y)
            for (uint64_t z = mn[2]; z <= mx[2]; ++z) {
                uint64_t cell[3] = {x, y, z};
                uint64_t ci = cell_idx(g, x, y, z);
                float d = cell_distance(g, verts, tris, t, cell, sign);
                if (locks) {
                    /* generate/grid.rs:445-454: unlocked read, then re-check under the lock */
                    float seen;
                    uint32_t bits = __atomic_load_n((uint32_t*)&pre_d[ci], __ATOMIC_RELAXED);
                    memcpy(&seen, &bits, 4);
                    if (orc_compare_distances(d, seen) == -1) {
                        cell_lock(locks, ci);
                        if (orc_compare_distances(d, pre_d[ci]) == -1) {
                            pre_d[ci] = d;
                            pre_t[ci] = t;
                            updates++;
                        }
                        cell_unlock(locks, ci);
                    }
                } else if (orc_compare_distances(d, pre_d[ci]) == -1) {
                    pre_d[ci] = d;
                    pre_t[ci] = t;
                    updates++;
                }
            }
    return updates;
}

/* generate/grid.rs:495-558 for one heap */
static int propagate(const ogrid* g, const float* verts, const uint32_t* tris, oheap* h, float* dist_arr,
                     uint32_t* dist_tri, int sign, uint8_t* locks, uint64_t* pops) {
    ostate s;
    uint64_t np = 0;
    while (heap_pop(h, &s)) {
        np++;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    int64_t x = (int64_t)s.cell[0] + dx, y = (int64_t)s.cell[1] + dy, z = (int64_t)s.cell[2] + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= (int64_t)g->n[0] || y >= (int64_t)g->n[1] ||
                        z >= (int64_t)g->n[2])
                        continue;
                    uint64_t cell[3] = {(uint64_t)x, (uint64_t)y, (uint64_t)z};
                    uint64_t ci = cell_idx(g, cell[0], cell[1], cell[2]);
                    float d = cell_distance(g, verts, tris, s.tri, cell, sign);
                    int improved = 0;
                    if (locks) cell_lock(locks, ci);
                    if (orc_compare_distances(d, dist_arr[ci]) == -1) {
                        dist_arr[ci] = d;
                        if (dist_tri) dist_tri[ci] = s.tri; /* the triangle the pushed state carries */
                        improved = 1;
                    }
                    if (locks) cell_unlock(locks, ci);
                    if (improved) {
                        ostate ns = {d, {(uint32_t)x, (uint32_t)y, (uint32_t)z}, s.tri};
                        if (heap_push(h, ns)) return -1;
                    }
                }
    }
    *pops += np;
    return 0;
}

/* ---- the frontier schedule (see sdf_oracle.h): packed keys, atomic min ---- */

static const uint64_t KEY_EMPTY = UINT64_MAX;

/* A candidate is stored only where compare_distances(d, f32::MAX) is Less
 * (grid.rs:121, 446): |d| below f32::MAX and not NaN, and for a negative d not
 * within 2 ulps of f32::MAX either (those approximately equal f32::MAX, which
 * is positive, so they come after it). */
static inline int key_valid(float d) {
    float a = fabsf(d);
    uint32_t bits;
    memcpy(&bits, &a, 4);
    return a < FLT_MAX && (!sign_negative(d) || bits < 0x7f7ffffdu);
}
/* Raycast: (bits(d) << 32 | triangle); Normal: (bits(|d|) << 32 | negative << 31 | triangle) */
static inline uint64_t make_key(float d, uint32_t t, int sign) {
    float a = sign == ORC_SIGN_RAYCAST ? d : fabsf(d);
    uint32_t bits;
    memcpy(&bits, &a, 4);
    uint64_t k = ((uint64_t)bits << 32) | t;
    if (sign == ORC_SIGN_NORMAL) k |= (uint64_t)sign_negative(d) << 31;
    return k;
}
static inline float key_value(uint64_t k, int sign) {
    if (k == KEY_EMPTY) return FLT_MAX;
    uint32_t bits = (uint32_t)(k >> 32);
    float a;
    memcpy(&a, &bits, 4);
    return (sign == ORC_SIGN_NORMAL && ((k >> 31) & 1u)) ? -a : a;
}
static inline uint32_t key_tri(uint64_t k, int sign) {
    return sign == ORC_SIGN_NORMAL ? (uint32_t)(k & 0x7fffffffu) : (uint32_t)k;
}
/* Whether key a comes strictly before key b in the order the reference keeps
 * (lib.rs:315-332 compare_distances, "Less" replaces the stored value), with
 * the lower triangle first among equal distances. Raycast values are never
 * negative, so compare_distances is plain < on them and the packed key's
 * integer order is that order. In Normal mode compare_distances orders
 * values by |d| except that a positive value comes before a negative one it
 * approximately equals (float-cmp's approx_eq, ulps 2, epsilon 1e-6); it is a
 * strict weak order (the positive values it places before a negative |d| = y
 * are those up to a threshold that grows with y), so this is a strict total
 * order and the minimum under it does not depend on the order the candidates
 * arrive in. An empty cell (f32::MAX) comes after every valid key. */
static int g_fault;
static inline int key_before(uint64_t a, uint64_t b, int sign) {
    if (g_fault == ORC_FAULT_PACKED_ORDER) return a < b;
    if (sign == ORC_SIGN_NORMAL && a != KEY_EMPTY && b != KEY_EMPTY && ((a ^ b) >> 31 & 1u)) {
        uint32_t ab = (uint32_t)(a >> 32), bb = (uint32_t)(b >> 32);
        float fa, fb;
        memcpy(&fa, &ab, 4);
        memcpy(&fb, &bb, 4);
        if (orc_approx_eq_f32(fa, fb, 2, 1e-6f)) return (int)((b >> 31) & 1u); /* the positive one first */
    }
    return a < b;
}
/* Lowers *p to v when v comes before it (key_before); returns 1 if it did. */
static inline int key_lower(uint64_t* p, uint64_t v, int sign) {
    uint64_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (key_before(v, cur, sign))
        if (__atomic_compare_exchange_n(p, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return 1;
    return 0;
}

/* Test-only fault injection into the frontier schedule (orc_set_fault), so the
 * tests can show that the fixpoint check catches a schedule that is wrong. */
void orc_set_fault(int kind) { g_fault = kind; }

/* ---------------------------------------------- supporting-plane skip check */
static int g_bound_mode;
static uint64_t g_bound_counts[3];
void orc_set_bound_check(int mode) {
    g_bound_mode = mode;
    g_bound_counts[0] = g_bound_counts[1] = g_bound_counts[2] = 0;
}
void orc_bound_check_counts(uint64_t* counts) {
    for (int i = 0; i < 3; ++i) counts[i] = g_bound_counts[i];
}
/* dir_pack: the vector over its largest component, three signed 10-bit fields */
static uint32_t bound_dir_pack(float dx, float dy, float dz) {
    float m = fmaxf(fmaxf(fabsf(dx), fabsf(dy)), fabsf(dz));
    if (!(m > 0.0f) || !(m <= 3.4e38f)) return 0u;
    float sc = 511.0f / m;
    int qx = (int)rintf(dx * sc), qy = (int)rintf(dy * sc), qz = (int)rintf(dz * sc);
    return ((uint32_t)qx & 1023u) | (((uint32_t)qy & 1023u) << 10) | (((uint32_t)qz & 1023u) << 20);
}
/* plane_bound: u from the word, normalised; no bound for a zero word or a
 * sliver; g = u.c - max_i u.v_i - (2^-16 largest coordinate + 2^-12 longest edge) */
static void plane_bound_c(uint32_t w, v3 c, v3 A, v3 B, v3 C, float max_size, float* u4) {
    u4[0] = u4[1] = u4[2] = 0.0f;
    u4[3] = -INFINITY;
    float ux = (float)(((int32_t)(w << 22)) >> 22), uy = (float)(((int32_t)(w << 12)) >> 22),
          uz = (float)(((int32_t)(w << 2)) >> 22);
    float l2 = ux * ux + uy * uy + uz * uz;
    v3 ab = vsub(B, A), ac = vsub(C, A), bc = vsub(C, B), n = vcross(ab, ac);
    float ab2 = vdot(ab, ab), ac2 = vdot(ac, ac), bc2 = vdot(bc, bc), n2 = vdot(n, n);
    int degen = veq(A, B) || veq(B, C) || veq(A, C);
    if (!(l2 > 0.0f) || (!degen && !(n2 >= (ab2 * ac2) * (1.0f / 256.0f)))) return;
    float inv = 1.0f / sqrtf(l2);
    ux *= inv;
    uy *= inv;
    uz *= inv;
    v3 u = mk(ux, uy, uz);
    float h = fmaxf(vdot(u, A), fmaxf(vdot(u, B), vdot(u, C)));
    float sc = fmaxf(fmaxf(fabsf(c.x), fabsf(c.y)), fabsf(c.z));
    sc = fmaxf(sc, fmaxf(fmaxf(fabsf(A.x), fabsf(A.y)), fabsf(A.z)));
    sc = fmaxf(sc, fmaxf(fmaxf(fabsf(B.x), fabsf(B.y)), fabsf(B.z)));
    sc = fmaxf(sc, fmaxf(fmaxf(fabsf(C.x), fabsf(C.y)), fabsf(C.z)));
    sc = sc + 2.0f * max_size;  /* the farthest neighbour is at most two cells away */
    float e = sqrtf(fmaxf(ab2, fmaxf(ac2, bc2)));
    float margin = sc * (1.0f / 65536.0f) + e * (1.0f / 4096.0f);
    if (g_fault == ORC_FAULT_BOUND_MARGIN_SIGN) margin = -margin;
    float g = (vdot(u, c) - h) - margin;
    u4[0] = ux;
    u4[1] = uy;
    u4[2] = uz;
    u4[3] = g == g ? g : -INFINITY;
}

static void frontier_seed_triangle(const ogrid* g, const float* verts, const uint32_t* tris, uint32_t t, int sign,
                                   uint64_t* keys, int* nan_seen) {
    v3 lo, hi;
    bbox(TRI_A(t), TRI_B(t), TRI_C(t), &lo, &hi);
    uint64_t mn[3], mx[3];
    snap(g, lo, mn);
    snap(g, hi, mx);
    v3 mnf = cell_center(g, mn[0], mn[1], mn[2]);
    for (int i = 0; i < 3; ++i)
        if (mn[i] > 0 && get(mnf, i) > get(lo, i)) mn[i] -= 1;
    v3 mxf = cell_center(g, mx[0], mx[1], mx[2]);
    for (int i = 0; i < 3; ++i)
        if (mx[i] < g->n[i] - 1 && get(mxf, i) < get(hi, i)) mx[i] += 1;
    for (uint64_t x = mn[0]; x <= mx[0]; ++x)
        for (uint64_t y = mn[1]; y <= mx[1]; ++y)
            for (uint64_t z = mn[2]; z <= mx[2]; ++z) {
                uint64_t cell[3] = {x, y, z};
                float d = cell_distance(g, verts, tris, t, cell, sign);
                if (d != d) *nan_seen = 1;
                if (key_valid(d)) key_lower(&keys[cell_idx(g, x, y, z)], make_key(d, t, sign), sign);
            }
}

unsigned orc_grid_reach(float max_size, float abs_size) { return max_size / abs_size >= 1.5f ? 2u : 1u; }

/* Returns the number of rounds, or -1 on a NaN distance (the reference panics). */
static int64_t frontier_schedule(const ogrid* g, const float* verts, const uint32_t* tris, uint64_t nt, int sign,
                                 float* dist_arr, uint32_t* tri_out, float* seed_out, int n_threads,
                                 uint64_t* evals) {
    uint64_t ncells = g->n[0] * g->n[1] * g->n[2];
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (ncells + 1));
    uint8_t* changed = (uint8_t*)calloc(ncells + 1, 1);
    uint64_t* fr_cell = (uint64_t*)malloc(sizeof(uint64_t) * (ncells + 1));
    uint32_t* fr_tri = (uint32_t*)malloc(sizeof(uint32_t) * (ncells + 1));
    for (uint64_t i = 0; i < ncells; ++i) keys[i] = KEY_EMPTY;
    int nan_seen = 0;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 256) reduction(| : nan_seen)
    for (int64_t t = 0; t < (int64_t)nt; ++t) frontier_seed_triangle(g, verts, tris, (uint32_t)t, sign, keys, &nan_seen);
    if (seed_out)
        for (uint64_t i = 0; i < ncells; ++i) seed_out[i] = key_value(keys[i], sign);
    for (uint64_t i = 0; i < ncells; ++i) changed[i] = keys[i] != KEY_EMPTY;
    const float max_size = fmaxf(fmaxf(fabsf(g->size.x), fabsf(g->size.y)), fabsf(g->size.z));
    const float band = ORC_GRID_BAND_CELLS * max_size;
    const float far = ORC_GRID_FAR_CELLS * max_size;
    const unsigned reach[3] = {orc_grid_reach(max_size, fabsf(g->size.x)), orc_grid_reach(max_size, fabsf(g->size.y)),
                               orc_grid_reach(max_size, fabsf(g->size.z))};
    const uint64_t release = ncells / ORC_GRID_RELEASE_DIV > ORC_GRID_RELEASE_MIN ? ncells / ORC_GRID_RELEASE_DIV
                                                                                  : ORC_GRID_RELEASE_MIN;
    uint64_t prev_taken = UINT64_MAX; /* round 0 keeps the band */
    int released = 0;                 /* every later round takes every flagged cell */
    int thin_seen = 0;                /* a thin round has happened */
    int64_t rounds = 0;
    uint64_t n_evals = 0;
    for (;;) {
        /* smallest |distance| among the flagged cells (keys order as these bits) */
        uint32_t m = UINT32_MAX;
        for (uint64_t i = 0; i < ncells; ++i)
            if (changed[i] && (uint32_t)(keys[i] >> 32) < m) m = (uint32_t)(keys[i] >> 32);
        if (m == UINT32_MAX) break;
        float mf;
        memcpy(&mf, &m, 4);
        float limit = mf + band;
        if (prev_taken < release) {
            /* a thin round in the far field, or the second one */
            if (mf > far || (thin_seen && g_fault != ORC_FAULT_NO_SECOND_RELEASE)) released = 1;
            thin_seen = 1;
        }
        /* a NaN band takes every flagged cell */
        if (!(limit >= mf) || prev_taken < release || released) limit = INFINITY;
        uint64_t nf = 0;
        for (uint64_t i = 0; i < ncells; ++i) {
            if (!changed[i]) continue;
            uint32_t b = (uint32_t)(keys[i] >> 32);
            float v;
            memcpy(&v, &b, 4);
            if (!(v <= limit)) {
                if (g_fault == ORC_FAULT_DROP_DEFERRED) changed[i] = 0;
                continue;
            }
            changed[i] = 0;
            fr_cell[nf] = i;
            fr_tri[nf] = key_tri(keys[i], sign);
            nf++;
        }
        rounds++;
        prev_taken = nf;
        uint64_t bc_checked = 0, bc_skipped = 0, bc_wrong = 0;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 256) reduction(| : nan_seen) \
    reduction(+ : n_evals, bc_checked, bc_skipped, bc_wrong)
        for (int64_t f = 0; f < (int64_t)nf; ++f) {
            uint64_t ci = fr_cell[f];
            uint32_t t = fr_tri[f];
            int64_t cz = (int64_t)(ci % g->n[2]), cy = (int64_t)((ci / g->n[2]) % g->n[1]),
                    cx = (int64_t)(ci / (g->n[1] * g->n[2]));
            float bound[4] = {0.0f, 0.0f, 0.0f, -INFINITY};
            if (g_bound_mode) {
                v3 c = cell_center(g, (uint64_t)cx, (uint64_t)cy, (uint64_t)cz);
                v3 A = TRI_A(t), B = TRI_B(t), C = TRI_C(t);
                uint32_t w;
                if (g_bound_mode == 1) {
                    v3 q = closest_triangle(c, A, B, C);
                    w = bound_dir_pack(c.x - q.x, c.y - q.y, c.z - q.z);
                } else {
                    uint64_t hsh = (ci + 0x9E3779B97F4A7C15ull * (uint64_t)(f + 1)) * 0xBF58476D1CE4E5B9ull;
                    w = (uint32_t)(hsh >> 32);
                }
                plane_bound_c(w, c, A, B, C, max_size, bound);
            }
            for (int nb = 0; nb < 33; ++nb) {
                        int dx, dy, dz;
                        if (nb < 27) {
                            dx = nb / 9 - 1;
                            dy = (nb / 3) % 3 - 1;
                            dz = nb % 3 - 1;
                            if (!dx && !dy && !dz) continue; /* the cell itself can never improve */
                        } else { /* 27 + 2a, 28 + 2a... : two steps along axis a, both ways, when its reach is 2 */
                            int a = (nb - 27) / 2, sgn = ((nb - 27) & 1) ? -1 : 1;
                            if (reach[a] < 2) continue;
                            dx = a == 0 ? 2 * sgn : 0;
                            dy = a == 1 ? 2 * sgn : 0;
                            dz = a == 2 ? 2 * sgn : 0;
                        }
                        if (g_fault == ORC_FAULT_SKIP_DIRECTION && dx > 0 && !dy && !dz) continue;
                        int64_t x = cx + dx, y = cy + dy, z = cz + dz;
                        if (x < 0 || y < 0 || z < 0 || x >= (int64_t)g->n[0] || y >= (int64_t)g->n[1] ||
                            z >= (int64_t)g->n[2])
                            continue;
                        uint64_t cell[3] = {(uint64_t)x, (uint64_t)y, (uint64_t)z};
                        uint64_t ni = cell_idx(g, cell[0], cell[1], cell[2]);
                        int skip = 0;
                        uint64_t cur = 0;
                        if (g_bound_mode) {
                            cur = __atomic_load_n(&keys[ni], __ATOMIC_RELAXED);
                            if (cur != KEY_EMPTY && key_tri(cur, sign) != t) {
                                bc_checked++;
                                uint32_t cb = (uint32_t)(cur >> 32);
                                if (sign == ORC_SIGN_NORMAL && ((cur >> 31) & 1ull)) {
                                    float cf, r;
                                    uint32_t rb;
                                    memcpy(&cf, &cb, 4);
                                    r = cf + 2e-6f;
                                    memcpy(&rb, &r, 4);
                                    cb = rb > cb + 2u ? rb : cb + 2u;
                                }
                                float thr;
                                memcpy(&thr, &cb, 4);
                                float lb = bound[3] + (bound[0] * ((float)dx * g->size.x) + bound[1] * ((float)dy * g->size.y) +
                                                       bound[2] * ((float)dz * g->size.z));
                                if (lb > thr) {
                                    skip = 1;
                                    bc_skipped++;
                                }
                            }
                        }
                        float d = cell_distance(g, verts, tris, t, cell, sign);
                        n_evals++;
                        if (d != d) nan_seen = 1;
                        if (skip && key_valid(d) && key_before(make_key(d, t, sign), cur, sign)) bc_wrong++;
                        if (key_valid(d) && key_lower(&keys[ni], make_key(d, t, sign), sign))
                            __atomic_store_n(&changed[ni], 1, __ATOMIC_RELAXED);
                    }
        }
        g_bound_counts[0] += bc_checked;
        g_bound_counts[1] += bc_skipped;
        g_bound_counts[2] += bc_wrong;
    }
    for (uint64_t i = 0; i < ncells; ++i) dist_arr[i] = key_value(keys[i], sign);
    if (tri_out)
        for (uint64_t i = 0; i < ncells; ++i) tri_out[i] = keys[i] == KEY_EMPTY ? UINT32_MAX : key_tri(keys[i], sign);
    free(keys);
    free(changed);
    free(fr_cell);
    free(fr_tri);
    if (evals) *evals = n_evals;
    return nan_seen ? -1 : rounds;
}

typedef struct {
    const ogrid* g;
    const float* verts;
    const uint32_t* tris;
    oheap* heap;
    float* dist_arr;
    int sign;
    uint8_t* locks;
    uint64_t pops;
    int err;
} prop_job;

/* generate/grid.rs:568-642 with the rays of generate/grid.rs:648-684 */
static void grid_raycast_sign(const ogrid* gp, const float* verts, const uint32_t* tris, uint64_t nt, float* dist_arr,
                              int n_threads, orc_grid_stats* lstp) {
    ogrid g = *gp;
    uint64_t ncells = g.n[0] * g.n[1] * g.n[2];
    int threaded = n_threads > 1;
    int n_heaps = n_threads;
    orc_grid_stats lst = *lstp;
    obvh bvh;
    bvh_build(&bvh, verts, tris, nt);
    uint32_t* inter = (uint32_t*)calloc(3 * ncells, sizeof(uint32_t));
    uint64_t nrays_axis[3] = {g.n[1] * g.n[2], g.n[0] * g.n[2], g.n[0] * g.n[1]};
    uint64_t nrays = nrays_axis[0] + nrays_axis[1] + nrays_axis[2];
    uint64_t cand = 0;
    int nthr = threaded ? n_heaps : 1;
#pragma omp parallel for num_threads(nthr) schedule(dynamic, 64) reduction(+ : cand)
    for (int64_t r = 0; r < (int64_t)nrays; ++r) {
        int axis;
        uint64_t start[3];
        uint64_t rr = (uint64_t)r;
        if (rr < nrays_axis[0]) {
            axis = 0;
            start[0] = 0;
            start[1] = rr / g.n[2];
            start[2] = rr % g.n[2];
        } else if (rr < nrays_axis[0] + nrays_axis[1]) {
            rr -= nrays_axis[0];
            axis = 1;
            start[0] = rr / g.n[2];
            start[1] = 0;
            start[2] = rr % g.n[2];
        } else {
            rr -= nrays_axis[0] + nrays_axis[1];
            axis = 2;
            start[0] = rr / g.n[1];
            start[1] = rr % g.n[1];
            start[2] = 0;
        }
        v3 o = cell_center(&g, start[0], start[1], start[2]);
        float ts[4096];
        uint64_t c0 = 0;
        int nh = bvh_ray_hits(&bvh, verts, tris, o, axis, ts, 4096, &c0);
        cand += c0;
        if (nh > 4096) nh = 4096; /* never reached by the test meshes */
        for (int hI = 0; hI < nh; ++hI) {
            uint64_t kk = f2usize(floorf(ts[hI] / get(g.size, axis)));
            if (kk > g.n[axis] - 1) kk = g.n[axis] - 1;
            uint64_t cell[3] = {start[0], start[1], start[2]};
            for (uint64_t i = 0; i <= kk; ++i) {
                cell[axis] = i;
                uint64_t ci = cell_idx(&g, cell[0], cell[1], cell[2]);
                __atomic_fetch_add(&inter[3 * ci + axis], 1u, __ATOMIC_RELAXED);
            }
        }
    }
    lst.rays = nrays;
    lst.ray_candidates = cand;
    for (uint64_t i = 0; i < ncells; ++i) {
        int odd = (int)(inter[3 * i] % 2) + (int)(inter[3 * i + 1] % 2) + (int)(inter[3 * i + 2] % 2);
        if (odd >= 2) dist_arr[i] = -dist_arr[i];
    }
    free(inter);
    bvh_free(&bvh);
    *lstp = lst;
}

/* generate/grid.rs:265-378 */
int orc_generate_grid_sdf(const float* verts, uint64_t nv, const uint32_t* idx, uint64_t ni, int topo,
                          const float* first_cell, const float* cell_size, const uint64_t* cell_count, int sign,
                          float* out, int n_heaps, int schedule, float* seed_out, orc_grid_stats* stats) {
    return orc_generate_grid_sdf_ex(verts, nv, idx, ni, topo, first_cell, cell_size, cell_count, sign, out, NULL,
                                    n_heaps, schedule, seed_out, stats);
}

int orc_generate_grid_sdf_ex(const float* verts, uint64_t nv, const uint32_t* idx, uint64_t ni, int topo,
                             const float* first_cell, const float* cell_size, const uint64_t* cell_count, int sign,
                             float* out, uint32_t* tri_out, int n_heaps, int schedule, float* seed_out,
                             orc_grid_stats* stats) {
    if (n_heaps <= 0) n_heaps = 1;
    if (sign != ORC_SIGN_RAYCAST && sign != ORC_SIGN_NORMAL) return ORC_ERR_INVALID_ARGUMENT;
    orc_grid_stats lst;
    memset(&lst, 0, sizeof(lst));
    uint32_t* tris = NULL;
    uint64_t nt = 0;
    int s = decode_alloc(nv, idx, ni, topo, &tris, &nt);
    if (s) return s;
    ogrid g = {ld(first_cell), ld(cell_size), {cell_count[0], cell_count[1], cell_count[2]}};
    uint64_t ncells = g.n[0] * g.n[1] * g.n[2];
    if (nt > 0 && (g.n[0] == 0 || g.n[1] == 0 || g.n[2] == 0)) {
        /* snap_point_to_grid's clamp(0, n - 1) panics (grid.rs:311-315) */
        free(tris);
        return ORC_ERR_INVALID_GRID;
    }
    if (schedule == ORC_GRID_EXACT) {
        float* dist_arr = (float*)malloc(sizeof(float) * (ncells + 1));
        int st2 = ORC_OK;
        for (uint64_t i = 0; i < ncells; ++i) {
            dist_arr[i] = FLT_MAX; /* grid.rs:137-143, kept without triangles */
            if (tri_out) tri_out[i] = UINT32_MAX;
        }
        if (nt > 0 && ncells > 0 && sign == ORC_SIGN_RAYCAST) {
            obvh b;
            bvh_build(&b, verts, tris, nt);
            int nan_hit = 0;
#pragma omp parallel for num_threads(n_heaps) schedule(dynamic, 256) reduction(| : nan_hit)
            for (int64_t i = 0; i < (int64_t)ncells; ++i) {
                uint64_t z = (uint64_t)i % g.n[2], y = ((uint64_t)i / g.n[2]) % g.n[1], x = (uint64_t)i / (g.n[1] * g.n[2]);
                float bd;
                uint32_t bt;
                bvh_nearest(&b, verts, tris, cell_center(&g, x, y, z), &bd, &bt);
                if (bd != bd) nan_hit = 1;
                dist_arr[i] = sqrtf(bd); /* point.rs:114-121 length of the nearest difference */
                if (tri_out) tri_out[i] = bt;
            }
            bvh_free(&b);
            if (nan_hit) st2 = ORC_ERR_NAN_DISTANCE;
            else grid_raycast_sign(&g, verts, tris, nt, dist_arr, n_heaps, &lst);
        } else if (nt > 0 && ncells > 0) {
            float* centres = (float*)malloc(sizeof(float) * 3 * ncells);
            for (uint64_t i = 0; i < ncells; ++i) {
                uint64_t z = i % g.n[2], y = (i / g.n[2]) % g.n[1], x = i / (g.n[1] * g.n[2]);
                st(centres + 3 * i, cell_center(&g, x, y, z));
            }
            st2 = orc_generate_sdf(verts, nv, idx, ni, topo, centres, ncells, ORC_ACCEL_BVH, ORC_SIGN_NORMAL, dist_arr,
                                   n_heaps, NULL);
            free(centres);
        }
        if (st2 == ORC_OK) memcpy(out, dist_arr, sizeof(float) * ncells);
        free(dist_arr);
        free(tris);
        if (stats) *stats = lst;
        return st2;
    }
    if (schedule == ORC_GRID_FRONTIER) {
        float* dist_arr = (float*)malloc(sizeof(float) * (ncells + 1));
        uint64_t evals = 0;
        int64_t rounds = frontier_schedule(&g, verts, tris, nt, sign, dist_arr, tri_out, seed_out, n_heaps, &evals);
        lst.rounds = rounds < 0 ? 0 : (uint64_t)rounds;
        lst.evals = evals;
        int st2 = rounds < 0 ? ORC_ERR_NAN_DISTANCE : ORC_OK;
        if (st2 == ORC_OK && sign == ORC_SIGN_RAYCAST && nt > 0 && ncells > 0)
            grid_raycast_sign(&g, verts, tris, nt, dist_arr, n_heaps, &lst);
        if (st2 == ORC_OK) memcpy(out, dist_arr, sizeof(float) * ncells);
        free(dist_arr);
        free(tris);
        if (stats) *stats = lst;
        return st2;
    }
    int threaded = schedule == ORC_GRID_THREADED;
    float* pre_d = (float*)malloc(sizeof(float) * (ncells + 1));
    uint32_t* pre_t = (uint32_t*)malloc(sizeof(uint32_t) * (ncells + 1));
    float* dist_arr = (float*)malloc(sizeof(float) * (ncells + 1));
    uint8_t* locks = threaded ? (uint8_t*)calloc(ncells + 1, 1) : NULL;
    if (!pre_d || !pre_t || !dist_arr || (threaded && !locks)) {
        free(pre_d);
        free(pre_t);
        free(dist_arr);
        free(locks);
        free(tris);
        return ORC_ERR_INVALID_ARGUMENT;
    }
    for (uint64_t i = 0; i < ncells; ++i) {
        if (tri_out) tri_out[i] = UINT32_MAX; /* no value */
        pre_d[i] = FLT_MAX; /* grid.rs:118-124 */
        pre_t[i] = 0;
        dist_arr[i] = FLT_MAX; /* grid.rs:137-143 */
    }

    /* seeding */
    if (threaded) {
        uint64_t upd = 0;
#pragma omp parallel for num_threads(n_heaps) schedule(dynamic, 256) reduction(+ : upd)
        for (int64_t t = 0; t < (int64_t)nt; ++t) upd += seed_triangle(&g, verts, tris, (uint32_t)t, sign, pre_d, pre_t, locks);
        lst.seed_updates = upd;
    } else {
        for (uint64_t t = 0; t < nt; ++t)
            lst.seed_updates += seed_triangle(&g, verts, tris, (uint32_t)t, sign, pre_d, pre_t, NULL);
    }
    if (seed_out) memcpy(seed_out, pre_d, sizeof(float) * ncells);

    /* generate/grid.rs:464-490 */
    uint64_t nseed = 0;
    for (uint64_t i = 0; i < ncells; ++i)
        if (pre_d[i] < FLT_MAX) nseed++;
    ostate* states = (ostate*)malloc(sizeof(ostate) * (nseed + 1));
    uint64_t k = 0;
    for (uint64_t i = 0; i < ncells; ++i)
        if (pre_d[i] < FLT_MAX) {
            uint64_t z = i % g.n[2], y = (i / g.n[2]) % g.n[1], x = i / (g.n[1] * g.n[2]);
            dist_arr[i] = pre_d[i];
            if (tri_out) tri_out[i] = pre_t[i];
            ostate st0 = {pre_d[i], {(uint32_t)x, (uint32_t)y, (uint32_t)z}, pre_t[i]};
            states[k++] = st0;
        }
    g_sort_tris = tris;
    qsort(states, nseed, sizeof(ostate), qsort_state_cmp);
    lst.seeded_cells = nseed;

    /* generate/grid.rs:318-339: round-robin split, heaps taken from the back */
    prop_job* jobs = (prop_job*)calloc((size_t)n_heaps, sizeof(prop_job));
    oheap* heaps = (oheap*)calloc((size_t)n_heaps, sizeof(oheap));
    for (int hI = 0; hI < n_heaps; ++hI) {
        heaps[hI].tris = tris;
        heaps[hI].cap = nseed / (uint64_t)n_heaps + 1024;
        heaps[hI].data = (ostate*)malloc(sizeof(ostate) * heaps[hI].cap);
    }
    for (uint64_t i = 0; i < nseed; ++i) {
        oheap* h = &heaps[i % (uint64_t)n_heaps];
        h->data[h->len++] = states[i];
    }
    free(states);
    for (int hI = 0; hI < n_heaps; ++hI) {
        heap_rebuild(&heaps[hI]);
        prop_job j = {&g, verts, tris, &heaps[hI], dist_arr, sign, locks, 0, 0};
        jobs[hI] = j;
    }
    int err = 0;
    if (threaded) {
#pragma omp parallel for num_threads(n_heaps) schedule(static, 1)
        for (int hI = n_heaps - 1; hI >= 0; --hI)
            jobs[hI].err = propagate(&g, verts, tris, jobs[hI].heap, dist_arr, tri_out, sign, locks, &jobs[hI].pops);
    } else {
        for (int hI = n_heaps - 1; hI >= 0; --hI)
            jobs[hI].err = propagate(&g, verts, tris, jobs[hI].heap, dist_arr, tri_out, sign, NULL, &jobs[hI].pops);
    }
    for (int hI = 0; hI < n_heaps; ++hI) {
        lst.pops += jobs[hI].pops;
        err |= jobs[hI].err;
        free(heaps[hI].data);
    }
    free(heaps);
    free(jobs);

    if (!err && sign == ORC_SIGN_RAYCAST && nt > 0 && ncells > 0)
        grid_raycast_sign(&g, verts, tris, nt, dist_arr, threaded ? n_heaps : 1, &lst);
    if (!err) memcpy(out, dist_arr, sizeof(float) * ncells);
    free(pre_d);
    free(pre_t);
    free(dist_arr);
    free(locks);
    free(tris);
    if (stats) *stats = lst;
    return err ? ORC_ERR_INVALID_ARGUMENT : ORC_OK;
}

/* ----------------------------------------------------------------- helpers */

void orc_nearest_bruteforce(const float* verts, const uint32_t* tris, uint64_t nt, const float* points,
                            uint64_t np, float* d2_out, uint32_t* tri_out, int n_threads) {
    if (n_threads <= 0) n_threads = 1;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)np; ++i) {
        v3 p = ld(points + 3 * i);
        float bd = INFINITY;
        uint32_t bt = UINT32_MAX;
        for (uint64_t t = 0; t < nt; ++t) {
            float d2 = tri_distance2(p, TRI_A(t), TRI_B(t), TRI_C(t));
            if (d2 < bd) {
                bd = d2;
                bt = (uint32_t)t;
            }
        }
        d2_out[i] = bd;
        if (tri_out) tri_out[i] = bt;
    }
}

/* orc_nearest_bruteforce's result through the oracle's BVH (conservative
 * pruning, so the same minimum FP32 d^2 and the same lowest triangle), for
 * point sets where brute force would take too long. */
void orc_nearest_bvh(const float* verts, const uint32_t* tris, uint64_t nt, const float* points, uint64_t np,
                     float* d2_out, uint32_t* tri_out, int n_threads) {
    if (n_threads <= 0) n_threads = 1;
    obvh b;
    bvh_build(&b, verts, tris, nt);
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)np; ++i) {
        float bd;
        uint32_t bt;
        bvh_nearest(&b, verts, tris, ld(points + 3 * i), &bd, &bt);
        d2_out[i] = bd;
        if (tri_out) tri_out[i] = bt;
    }
    bvh_free(&b);
}

/* The state every schedule of the reference's propagation reaches
 * (generate/grid.rs:495-558): when the last state a cell c was pushed with is
 * popped, its triangle t(c) is offered to every neighbour n, and afterwards
 * n's value only moves earlier in compare_distances' order, so at the end
 *   (1) value(c) is the distance from c's centre to t(c), bit for bit
 *       (|value| for Raycast, whose sign comes from the raycasts),
 *   (2) no neighbour n has d(centre(n), t(c)) strictly before value(n) under
 *       compare_distances (lib.rs:315-332; f32::MAX where n holds no value),
 * for every cell c holding a value (tri_of[c] != UINT32_MAX). counts[0] and
 * counts[1] receive the violations of (1) and (2), counts[2] the cells without
 * a triangle and counts[3] those of them whose value is not f32::MAX. */
int orc_grid_fixpoint_check(const float* verts, uint64_t nv, const uint32_t* idx, uint64_t ni, int topo,
                            const float* first_cell, const float* cell_size, const uint64_t* cell_count, int sign,
                            const float* values, const uint32_t* tri_of, int n_threads, uint64_t* counts) {
    if (n_threads <= 0) n_threads = 1;
    uint32_t* tris = NULL;
    uint64_t nt = 0;
    int s = decode_alloc(nv, idx, ni, topo, &tris, &nt);
    if (s) return s;
    ogrid g = {ld(first_cell), ld(cell_size), {cell_count[0], cell_count[1], cell_count[2]}};
    const uint64_t ncells = g.n[0] * g.n[1] * g.n[2];
    uint64_t bad_value = 0, not_fix = 0, no_tri = 0, no_tri_value = 0;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 1024) \
    reduction(+ : bad_value, not_fix, no_tri, no_tri_value)
    for (int64_t ii = 0; ii < (int64_t)ncells; ++ii) {
        const uint64_t i = (uint64_t)ii;
        const uint32_t t = tri_of[i];
        const float v = sign == ORC_SIGN_RAYCAST ? fabsf(values[i]) : values[i];
        if (t == UINT32_MAX) {
            no_tri++;
            if (!(v == FLT_MAX)) no_tri_value++;
            continue;
        }
        if (t >= nt) {
            bad_value++;
            continue;
        }
        const uint64_t cz = i % g.n[2], cy = (i / g.n[2]) % g.n[1], cx = i / (g.n[1] * g.n[2]);
        const uint64_t cell[3] = {cx, cy, cz};
        const float dc = cell_distance(&g, verts, tris, t, cell, sign);
        uint32_t b1, b2;
        memcpy(&b1, &dc, 4);
        memcpy(&b2, &v, 4);
        if (b1 != b2) bad_value++;
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    const int64_t x = (int64_t)cx + dx, y = (int64_t)cy + dy, z = (int64_t)cz + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= (int64_t)g.n[0] || y >= (int64_t)g.n[1] ||
                        z >= (int64_t)g.n[2])
                        continue;
                    const uint64_t nc[3] = {(uint64_t)x, (uint64_t)y, (uint64_t)z};
                    const uint64_t n = cell_idx(&g, nc[0], nc[1], nc[2]);
                    const float dn = cell_distance(&g, verts, tris, t, nc, sign);
                    const float vn = tri_of[n] == UINT32_MAX ? FLT_MAX
                                     : (sign == ORC_SIGN_RAYCAST ? fabsf(values[n]) : values[n]);
                    if (orc_compare_distances(dn, vn) == -1) not_fix++;
                }
    }
    counts[0] = bad_value;
    counts[1] = not_fix;
    counts[2] = no_tri;
    counts[3] = no_tri_value;
    free(tris);
    return ORC_OK;
}

typedef struct {
    double x, y, z;
} d3;
static inline d3 dmk(double x, double y, double z) {
    d3 r = {x, y, z};
    return r;
}
static inline d3 dsub(d3 a, d3 b) { return dmk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 dadd(d3 a, d3 b) { return dmk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 dmul(d3 a, double s) { return dmk(a.x * s, a.y * s, a.z * s); }
static inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 dld(const float* p) { return dmk(p[0], p[1], p[2]); }

/* Ericson, Real-Time Collision Detection 5.1.5, in double precision */
static d3 closest_fp64(d3 p, d3 a, d3 b, d3 c) {
    d3 ab = dsub(b, a), ac = dsub(c, a), ap = dsub(p, a);
    double d1 = ddot(ab, ap), d2 = ddot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return a;
    d3 bp = dsub(p, b);
    double d3_ = ddot(ab, bp), d4 = ddot(ac, bp);
    if (d3_ >= 0 && d4 <= d3_) return b;
    double vc = d1 * d4 - d3_ * d2;
    if (vc <= 0 && d1 >= 0 && d3_ <= 0) return dadd(a, dmul(ab, d1 / (d1 - d3_)));
    d3 cp = dsub(p, c);
    double d5 = ddot(ab, cp), d6 = ddot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return c;
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return dadd(a, dmul(ac, d2 / (d2 - d6)));
    double va = d3_ * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3_) >= 0 && (d5 - d6) >= 0)
        return dadd(b, dmul(dsub(c, b), (d4 - d3_) / ((d4 - d3_) + (d5 - d6))));
    double den = va + vb + vc;
    if (den == 0.0) {
        /* degenerate: nearest of the three edges */
        d3 cands[3];
        d3 e[3][2] = {{a, b}, {a, c}, {b, c}};
        double best = INFINITY;
        d3 bp2 = a;
        for (int i = 0; i < 3; ++i) {
            d3 s = dsub(e[i][1], e[i][0]);
            double m = ddot(s, s);
            double tt = m > 0 ? ddot(s, dsub(p, e[i][0])) / m : 0;
            tt = tt < 0 ? 0 : (tt > 1 ? 1 : tt);
            cands[i] = dadd(e[i][0], dmul(s, tt));
            d3 dd = dsub(p, cands[i]);
            if (ddot(dd, dd) < best) {
                best = ddot(dd, dd);
                bp2 = cands[i];
            }
        }
        return bp2;
    }
    double v = vb / den, w = vc / den;
    return dadd(a, dadd(dmul(ab, v), dmul(ac, w)));
}

void orc_fp64_distance(const float* verts, const uint32_t* tris, uint64_t nt, const float* points, uint64_t np,
                       double* out, int n_threads) {
    if (n_threads <= 0) n_threads = 1;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)np; ++i) {
        d3 p = dld(points + 3 * i);
        double best = INFINITY;
        for (uint64_t t = 0; t < nt; ++t) {
            d3 cp = closest_fp64(p, dld(verts + 3 * tris[3 * t]), dld(verts + 3 * tris[3 * t + 1]),
                                 dld(verts + 3 * tris[3 * t + 2]));
            d3 dd = dsub(p, cp);
            double d2 = ddot(dd, dd);
            if (d2 < best) best = d2;
        }
        out[i] = sqrt(best);
    }
}

void orc_ray_parity_bruteforce(const float* verts, const uint32_t* tris, uint64_t nt, const float* points,
                               uint64_t np, uint8_t* parity_out, int n_threads) {
    if (n_threads <= 0) n_threads = 1;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)np; ++i) {
        v3 p = ld(points + 3 * i);
        for (int axis = 0; axis < 3; ++axis) {
            uint32_t n = 0;
            for (uint64_t t = 0; t < nt; ++t) n += (uint32_t)ray_aligned(p, TRI_A(t), TRI_B(t), TRI_C(t), axis, NULL);
            parity_out[3 * i + axis] = (uint8_t)(n & 1u);
        }
    }
}

/* ---------------------------------------------------------- mesh_check.rs */

/* Debug formatting of f32 tells every non-NaN bit pattern apart (including
 * -0.0) and prints every NaN as "NaN". */
static int vertex_key_eq(const float* a, const float* b) {
    for (int k = 0; k < 3; ++k) {
        int an = a[k] != a[k], bn = b[k] != b[k];
        if (an || bn) {
            if (an != bn) return 0;
            continue;
        }
        if (memcmp(&a[k], &b[k], 4) != 0) return 0;
    }
    return 1;
}

/* generate/mesh_check.rs:24-81, as written: O(V^2) duplicate vertices (the
 * first vertex equal to an earlier one, :37-43), O(F^2) duplicate faces (the
 * first face equal to an earlier one as an ordered triple, :46-49), then the
 * edge loop of :52-78 that pops the last edge, fails on a remaining copy of it
 * or on a missing reverse edge, and removes every copy of the reverse.
 * Returns 0 (Ok), 1 (not a multiple of 3), 2 (duplicate vertex), 3 (duplicate
 * face), 4 (duplicate edge), 5 (edge without its reverse); *where_out gets the
 * vertex index (2), the face index (3) or the edge's index 3 * face + k for the
 * edge (face[k], face[(k + 1) % 3]) (4, 5). */
int orc_check_mesh_triangle_list_at(const float* verts, uint64_t nv, const uint32_t* faces, uint64_t nf,
                                    uint64_t* where_out) {
    *where_out = 0;
    if (nf % 3 != 0) return 1;
    for (uint64_t i = 0; i < nv; ++i)
        for (uint64_t j = 0; j < i; ++j)
            if (vertex_key_eq(verts + 3 * i, verts + 3 * j)) {
                *where_out = i;
                return 2;
            }
    uint64_t nt = nf / 3;
    for (uint64_t i = 0; i < nt; ++i)
        for (uint64_t j = 0; j < i; ++j)
            if (faces[3 * i] == faces[3 * j] && faces[3 * i + 1] == faces[3 * j + 1] &&
                faces[3 * i + 2] == faces[3 * j + 2]) {
                *where_out = i;
                return 3;
            }
    uint64_t ne = 3 * nt;
    uint32_t* e = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (ne + 1));
    uint64_t* origin = (uint64_t*)malloc(sizeof(uint64_t) * (ne + 1));
    for (uint64_t f = 0; f < nt; ++f)
        for (int k = 0; k < 3; ++k) {
            e[2 * (3 * f + k)] = faces[3 * f + k];
            e[2 * (3 * f + k) + 1] = faces[3 * f + (k + 1) % 3];
            origin[3 * f + k] = 3 * f + (uint64_t)k;
        }
    int result = 0;
    while (ne) {
        ne--;
        uint32_t a = e[2 * ne], b = e[2 * ne + 1];
        uint64_t same = 0;
        int has_opp = 0;
        for (uint64_t i = 0; i < ne; ++i) {
            if (e[2 * i] == a && e[2 * i + 1] == b) same++;
            if (e[2 * i] == b && e[2 * i + 1] == a) has_opp = 1;
        }
        if (same) {
            result = 4;
            *where_out = origin[ne];
            break;
        }
        if (!has_opp) {
            result = 5;
            *where_out = origin[ne];
            break;
        }
        uint64_t w = 0;
        for (uint64_t i = 0; i < ne; ++i)
            if (!(e[2 * i] == b && e[2 * i + 1] == a)) {
                e[2 * w] = e[2 * i];
                e[2 * w + 1] = e[2 * i + 1];
                origin[w] = origin[i];
                w++;
            }
        ne = w;
    }
    free(e);
    free(origin);
    return result;
}

int orc_check_mesh_triangle_list(const float* verts, uint64_t nv, const uint32_t* faces, uint64_t nf) {
    uint64_t where;
    return orc_check_mesh_triangle_list_at(verts, nv, faces, nf, &where);
}
