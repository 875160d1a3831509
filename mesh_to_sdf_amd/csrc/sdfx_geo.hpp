// sdfx_geo.hpp — FP32 point/triangle geometry evaluated exactly as the
// reference's Rust code rounds it (compile with -ffp-contract=off: no FMA
// contraction; IEEE division and sqrt). Citations are to mesh_to_sdf/src/.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace sdfx {

#define SDFX_HD __host__ __device__ __forceinline__

// One triangle, precomputed once per mesh. The edge vectors and the normal are
// the exact FP32 values the reference recomputes per query (b - a, c - a,
// c - b, cross(b - a, c - a)), so using them changes no rounding.
struct __attribute__((aligned(16))) TriRec {
    float ax, ay, az, bx;
    float by, bz, cx, cy;
    float cz, abx, aby, abz;
    float acx, acy, acz, bcx;
    float bcy, bcz, nx, ny;
    float nz;
    uint32_t degen;  // 0 generic, 1 a==b==c, 2 a==b (segment a-c), 3 b==c or a==c (segment a-b)
    uint32_t index;  // triangle index in the mesh's topology order
    uint32_t pad;
};
static_assert(sizeof(TriRec) == 96, "TriRec layout");

struct Box {
    float lo[3], hi[3];
};

SDFX_HD float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    // point.rs:100-102: ((x*x' + y*y') + z*z')
    return ax * bx + ay * by + az * bz;
}

// geo.rs:4-22: component min/max, then -/+ 1e-4 absolute padding.
SDFX_HD void tri_bbox(float ax, float ay, float az, float bx, float by, float bz, float cx, float cy, float cz,
                      Box& b) {
    const float eps = 0.0001f;
    b.lo[0] = fminf(ax, fminf(bx, cx)) - eps;
    b.lo[1] = fminf(ay, fminf(by, cy)) - eps;
    b.lo[2] = fminf(az, fminf(bz, cz)) - eps;
    b.hi[0] = fmaxf(ax, fmaxf(bx, cx)) + eps;
    b.hi[1] = fmaxf(ay, fmaxf(by, cy)) + eps;
    b.hi[2] = fmaxf(az, fmaxf(bz, cz)) + eps;
}

SDFX_HD void make_trirec(float ax, float ay, float az, float bx, float by, float bz, float cx, float cy, float cz,
                         uint32_t index, TriRec& t) {
    t.ax = ax; t.ay = ay; t.az = az;
    t.bx = bx; t.by = by; t.bz = bz;
    t.cx = cx; t.cy = cy; t.cz = cz;
    t.abx = bx - ax; t.aby = by - ay; t.abz = bz - az;
    t.acx = cx - ax; t.acy = cy - ay; t.acz = cz - az;
    t.bcx = cx - bx; t.bcy = cy - by; t.bcz = cz - bz;
    // point.rs:105-111 cross(ab, ac), geo.rs:60-64
    t.nx = t.aby * t.acz - t.abz * t.acy;
    t.ny = t.abz * t.acx - t.abx * t.acz;
    t.nz = t.abx * t.acy - t.aby * t.acx;
    bool eab = ax == bx && ay == by && az == bz;
    bool ebc = bx == cx && by == cy && bz == cz;
    bool eac = ax == cx && ay == cy && az == cz;
    // geo.rs:73-88 match order
    t.degen = (eab && ebc && eac) ? 1u : (eab ? 2u : ((ebc || eac) ? 3u : 0u));
    t.index = index;
    t.pad = 0;
}

// A triangle as the LBVH stores it: the vertices, the degenerate case and the
// topology index (48 bytes, three 16-byte loads). expand_leaf() rebuilds the
// full record in registers with make_trirec's expressions (each edge one FP32
// subtraction, the normal the same cross product), so every value, and every
// distance computed from them, is bit-identical to the stored TriRec's.
struct __attribute__((aligned(16))) LeafRec {
    float ax, ay, az, bx;
    float by, bz, cx, cy;
    float cz;
    uint32_t degen, index, pad;
};
static_assert(sizeof(LeafRec) == 48, "LeafRec layout");

SDFX_HD LeafRec make_leafrec(const TriRec& t) {
    return LeafRec{t.ax, t.ay, t.az, t.bx, t.by, t.bz, t.cx, t.cy, t.cz, t.degen, t.index, 0u};
}

SDFX_HD TriRec expand_leaf(const LeafRec& l) {
    TriRec t;
    make_trirec(l.ax, l.ay, l.az, l.bx, l.by, l.bz, l.cx, l.cy, l.cz, l.index, t);
    t.degen = l.degen;  // the same value make_trirec derives; kept to skip the comparisons
    return t;
}

// Closest point on the triangle to p (geo.rs:70-138), written as the vector
// p - closest so callers get dist/dist2/signed-direction directly. Vertex
// regions return p - a etc. (the same values as p.sub(a)).
SDFX_HD void closest_diff(float px, float py, float pz, const TriRec& t, float& dx, float& dy, float& dz,
                          float& cpx, float& cpy, float& cpz) {
    if (t.degen != 0u) {
        if (t.degen == 1u) {
            cpx = t.ax; cpy = t.ay; cpz = t.az;
        } else {
            // geo.rs:141-151 closest_point_segment(p, a, x), x = c (degen 2) or b (degen 3)
            float sx = t.degen == 2u ? t.acx : t.abx;
            float sy = t.degen == 2u ? t.acy : t.aby;
            float sz = t.degen == 2u ? t.acz : t.abz;
            float m = dot3(sx, sy, sz, sx, sy, sz);
            float apx = px - t.ax, apy = py - t.ay, apz = pz - t.az;
            float s = dot3(sx, sy, sz, apx, apy, apz) / m;
            s = s < 0.0f ? 0.0f : (s > 1.0f ? 1.0f : s);
            cpx = t.ax + sx * s; cpy = t.ay + sy * s; cpz = t.az + sz * s;
        }
        dx = px - cpx; dy = py - cpy; dz = pz - cpz;
        return;
    }
    float apx = px - t.ax, apy = py - t.ay, apz = pz - t.az;
    float d1 = dot3(t.abx, t.aby, t.abz, apx, apy, apz);
    float d2 = dot3(t.acx, t.acy, t.acz, apx, apy, apz);
    if (d1 <= 0.0f && d2 <= 0.0f) {
        cpx = t.ax; cpy = t.ay; cpz = t.az;
        dx = apx; dy = apy; dz = apz;
        return;
    }
    float bpx = px - t.bx, bpy = py - t.by, bpz = pz - t.bz;
    float d3 = dot3(t.abx, t.aby, t.abz, bpx, bpy, bpz);
    float d4 = dot3(t.acx, t.acy, t.acz, bpx, bpy, bpz);
    if (d3 >= 0.0f && d4 <= d3) {
        cpx = t.bx; cpy = t.by; cpz = t.bz;
        dx = bpx; dy = bpy; dz = bpz;
        return;
    }
    float cpx_ = px - t.cx, cpy_ = py - t.cy, cpz_ = pz - t.cz;
    float d5 = dot3(t.abx, t.aby, t.abz, cpx_, cpy_, cpz_);
    float d6 = dot3(t.acx, t.acy, t.acz, cpx_, cpy_, cpz_);
    if (d6 >= 0.0f && d5 <= d6) {
        cpx = t.cx; cpy = t.cy; cpz = t.cz;
        dx = cpx_; dy = cpy_; dz = cpz_;
        return;
    }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        float v = d1 / (d1 - d3);
        cpx = t.ax + t.abx * v; cpy = t.ay + t.aby * v; cpz = t.az + t.abz * v;
    } else {
        float vb = d5 * d2 - d1 * d6;
        if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
            float v = d2 / (d2 - d6);
            cpx = t.ax + t.acx * v; cpy = t.ay + t.acy * v; cpz = t.az + t.acz * v;
        } else {
            float va = d3 * d6 - d5 * d4;
            if (va <= 0.0f && d4 - d3 >= 0.0f && d5 - d6 >= 0.0f) {
                float v = (d4 - d3) / ((d4 - d3) + (d5 - d6));
                cpx = t.bx + t.bcx * v; cpy = t.by + t.bcy * v; cpz = t.bz + t.bcz * v;
            } else {
                float denom = 1.0f / (va + vb + vc);
                float v = vb * denom;
                float w = vc * denom;
                cpx = (t.ax + t.abx * v) + t.acx * w;
                cpy = (t.ay + t.aby * v) + t.acy * w;
                cpz = (t.az + t.abz * v) + t.acz * w;
            }
        }
    }
    dx = px - cpx; dy = py - cpy; dz = pz - cpz;
}

// The same closest point without per-lane branches on the Voronoi region:
// d1..d6 and va, vb, vc are always formed (the reference forms them lazily;
// their values do not depend on it), the region is the first of the
// reference's tests that holds (A, B, C, AB, AC, BC, else interior), and one
// IEEE division runs on the region's numerator and denominator:
//   AB: d1 / (d1 - d3)   AC: d2 / (d2 - d6)   BC: (d4 - d3) / ((d4 - d3) + (d5 - d6))
//   interior: 1 / (va + vb + vc)
// so every value, including the closest point, is bit-identical to
// closest_diff. Lanes of a wave no longer serialize on different regions.
SDFX_HD void closest_diff_flat(float px, float py, float pz, const TriRec& t, float& dx, float& dy, float& dz,
                               float& cpx, float& cpy, float& cpz) {
    if (t.degen != 0u) {
        closest_diff(px, py, pz, t, dx, dy, dz, cpx, cpy, cpz);
        return;
    }
    const float apx = px - t.ax, apy = py - t.ay, apz = pz - t.az;
    const float bpx = px - t.bx, bpy = py - t.by, bpz = pz - t.bz;
    const float cqx = px - t.cx, cqy = py - t.cy, cqz = pz - t.cz;
    const float d1 = dot3(t.abx, t.aby, t.abz, apx, apy, apz);
    const float d2 = dot3(t.acx, t.acy, t.acz, apx, apy, apz);
    const float d3 = dot3(t.abx, t.aby, t.abz, bpx, bpy, bpz);
    const float d4 = dot3(t.acx, t.acy, t.acz, bpx, bpy, bpz);
    const float d5 = dot3(t.abx, t.aby, t.abz, cqx, cqy, cqz);
    const float d6 = dot3(t.acx, t.acy, t.acz, cqx, cqy, cqz);
    const float vc = d1 * d4 - d3 * d2;
    const float vb = d5 * d2 - d1 * d6;
    const float va = d3 * d6 - d5 * d4;
    const float e43 = d4 - d3, e56 = d5 - d6;
    const bool rA = d1 <= 0.0f && d2 <= 0.0f;
    const bool rB = !rA && d3 >= 0.0f && d4 <= d3;
    const bool rC = !rA && !rB && d6 >= 0.0f && d5 <= d6;
    const bool vert = rA || rB || rC;
    const bool rAB = !vert && vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f;
    const bool rAC = !vert && !rAB && vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f;
    const bool rBC = !vert && !rAB && !rAC && va <= 0.0f && e43 >= 0.0f && e56 >= 0.0f;
    const bool inner = !vert && !rAB && !rAC && !rBC;
    const float num = rAB ? d1 : (rAC ? d2 : (rBC ? e43 : 1.0f));
    const float den = rAB ? (d1 - d3) : (rAC ? (d2 - d6) : (rBC ? (e43 + e56) : ((va + vb) + vc)));
    const float qq = num / den;
    const float s1 = inner ? vb * qq : qq;
    const float bx = rBC ? t.bx : t.ax, by = rBC ? t.by : t.ay, bz = rBC ? t.bz : t.az;
    const float ex = rAC ? t.acx : (rBC ? t.bcx : t.abx);
    const float ey = rAC ? t.acy : (rBC ? t.bcy : t.aby);
    const float ez = rAC ? t.acz : (rBC ? t.bcz : t.abz);
    float qx = bx + ex * s1, qy = by + ey * s1, qz = bz + ez * s1;
    if (inner) {
        const float w = vc * qq;
        qx = qx + t.acx * w;
        qy = qy + t.acy * w;
        qz = qz + t.acz * w;
    }
    cpx = rA ? t.ax : (rB ? t.bx : (rC ? t.cx : qx));
    cpy = rA ? t.ay : (rB ? t.by : (rC ? t.cy : qy));
    cpz = rA ? t.az : (rB ? t.bz : (rC ? t.cz : qz));
    // vertex regions: p - a etc. are the very values p.sub(vertex) gives
    dx = rA ? apx : (rB ? bpx : (rC ? cqx : px - qx));
    dy = rA ? apy : (rB ? bpy : (rC ? cqy : py - qy));
    dz = rA ? apz : (rB ? bpz : (rC ? cqz : pz - qz));
}

// geo.rs:33-37 point_triangle_distance2
SDFX_HD float tri_dist2(float px, float py, float pz, const TriRec& t) {
    float dx, dy, dz, cx, cy, cz;
    closest_diff_flat(px, py, pz, t, dx, dy, dz, cx, cy, cz);
    return dot3(dx, dy, dz, dx, dy, dz);
}

// geo.rs:26-30 point_triangle_distance (point.rs:114-121 length = sqrt(dot))
SDFX_HD float tri_dist(float px, float py, float pz, const TriRec& t) { return sqrtf(tri_dist2(px, py, pz, t)); }

// geo.rs:43-56 point_triangle_signed_distance: +d iff (p - cp) . n > 0, else -d
// (so a zero dot product gives -d, and d == 0 gives -0.0).
SDFX_HD float tri_signed_dist(float px, float py, float pz, const TriRec& t) {
    float dx, dy, dz, cx, cy, cz;
    closest_diff_flat(px, py, pz, t, dx, dy, dz, cx, cy, cz);
    float d = sqrtf(dot3(dx, dy, dz, dx, dy, dz));
    return dot3(dx, dy, dz, t.nx, t.ny, t.nz) > 0.0f ? d : -d;
}

// geo.rs:165-216 ray_triangle_intersection_aligned for ray direction +axis
// (0 = X, 1 = Y, 2 = Z). The (get_y, get_z, get_x) rotation of geo.rs:181-195
// is (y,z,x) for X, (z,x,y) for Y and (x,y,z) for Z. edge20 = a - c is -ac
// exactly (IEEE subtraction is sign-symmetric). Returns t > 0 or a negative value
// for no hit.
template <int AXIS>
SDFX_HD bool ray_hit(float ox, float oy, float oz, const TriRec& t, float* t_out) {
    constexpr int IY = AXIS == 0 ? 1 : (AXIS == 1 ? 2 : 0);
    constexpr int IZ = AXIS == 0 ? 2 : (AXIS == 1 ? 0 : 1);
    constexpr int IX = AXIS;
    const float o[3] = {ox, oy, oz};
    const float A[3] = {t.ax, t.ay, t.az};
    const float B[3] = {t.bx, t.by, t.bz};
    const float C[3] = {t.cx, t.cy, t.cz};
    const float e01[3] = {t.abx, t.aby, t.abz};
    const float e12[3] = {t.bcx, t.bcy, t.bcz};
    const float e20[3] = {-t.acx, -t.acy, -t.acz};
    float p0y = o[IY] - A[IY], p0z = o[IZ] - A[IZ];
    float p1y = o[IY] - B[IY], p1z = o[IZ] - B[IZ];
    float p2y = o[IY] - C[IY], p2z = o[IZ] - C[IZ];
    float w0 = p1z * e12[IY] - p1y * e12[IZ];
    float w1 = p2z * e20[IY] - p2y * e20[IZ];
    float w2 = p0z * e01[IY] - p0y * e01[IZ];
    if ((w0 < 0.0f && w1 < 0.0f && w2 < 0.0f) || (w0 > 0.0f && w1 > 0.0f && w2 > 0.0f)) {
        float p0x = o[IX] - A[IX], p1x = o[IX] - B[IX], p2x = o[IX] - C[IX];
        float tt = -(w0 * p0x + w2 * p2x + w1 * p1x) / (w0 + w1 + w2);
        if (tt > 0.0f) {
            if (t_out) *t_out = tt;
            return true;
        }
    }
    return false;
}

SDFX_HD bool ray_hit_axis(int axis, float ox, float oy, float oz, const TriRec& t, float* t_out) {
    return axis == 0 ? ray_hit<0>(ox, oy, oz, t, t_out)
                     : (axis == 1 ? ray_hit<1>(ox, oy, oz, t, t_out) : ray_hit<2>(ox, oy, oz, t, t_out));
}

SDFX_HD uint32_t fbits(float x) { return __builtin_bit_cast(uint32_t, x); }
SDFX_HD float bitsf(uint32_t u) { return __builtin_bit_cast(float, u); }

// float-cmp 0.9.0 approx_eq!(f32, a, b, ulps = 2, epsilon = 1e-6)
SDFX_HD bool approx_eq(float a, float b) {
    if (a == b) return true;
    if (fabsf(a - b) <= 1e-6f) return true;
    int32_t diff = (int32_t)(fbits(a) - fbits(b));
    int32_t ad = diff == INT32_MIN ? INT32_MAX : (diff < 0 ? -diff : diff);
    return ad <= 2;
}

// lib.rs:315-332 compare_distances: -1 Less, 0 Equal, 1 Greater, 2 NaN (the reference panics)
SDFX_HD int compare_distances(float a, float b) {
    float aa = fabsf(a), bb = fabsf(b);
    if (approx_eq(aa, bb)) {
        bool an = (fbits(a) >> 31) != 0u, bn = (fbits(b) >> 31) != 0u;
        if (an && !bn) return 1;
        if (!an && bn) return -1;
    }
    if (aa != aa || bb != bb) return 2;
    return aa < bb ? -1 : (aa > bb ? 1 : 0);
}

// Squared distance from p to a box (0 inside).
SDFX_HD float box_dist2(float px, float py, float pz, const float* lo, const float* hi) {
    float dx = fmaxf(fmaxf(lo[0] - px, px - hi[0]), 0.0f);
    float dy = fmaxf(fmaxf(lo[1] - py, py - hi[1]), 0.0f);
    float dz = fmaxf(fmaxf(lo[2] - pz, pz - hi[2]), 0.0f);
    return dx * dx + dy * dy + dz * dz;
}

// Pruning by box distance keeps a box when box_d2 * PRUNE_SLACK <= best_d2:
// the 2^-16 slack absorbs the rounding of the box distance, so a box is only
// dropped when every point in it is clearly farther than the best distance
// found, and the nearest triangle (minimum FP32 d^2) is never missed.
constexpr float PRUNE_SLACK = 1.0f - 1.0f / 65536.0f;

// Axes (bit k = +k ray) whose ray from p can meet the box: the two
// perpendicular coordinates inside, the box not entirely behind p.
SDFX_HD uint32_t ray_axes(const float* lo, const float* hi, float px, float py, float pz) {
    bool ix = lo[0] <= px && px <= hi[0];
    bool iy = lo[1] <= py && py <= hi[1];
    bool iz = lo[2] <= pz && pz <= hi[2];
    uint32_t m = 0;
    if (iy && iz && hi[0] >= px) m |= 1u;
    if (iz && ix && hi[1] >= py) m |= 2u;
    if (ix && iy && hi[2] >= pz) m |= 4u;
    return m;
}

}  // namespace sdfx
