// mlc.hpp -- multileaf collimator module (DYNVMLC-type, 60 leaf pairs).
//
// The MLC is a slab z in [zmin, zmin + zthick] of the head frame (z from the
// source).  Across the leaves (x-z plane) every leaf is a 14-vertex polygon
// with source-focused sides (cross_xz, leaf i at [16 i, 16 i + 14)); along the
// leaf travel (y-z plane) a leaf pair's ends are a 34-vertex profile
// (endnode: [0,1] the -rmax edge, [2,16] the bank A end, [17,31] the bank B
// end, [32,33] the +rmax edge) shifted to the field's leaf positions.  A
// particle in the module carries its leaf index (>= 0: inside leaf i's
// cross-section; -(i+1): in the air gap next to leaf i; OUTMLC: not yet
// located; PASSMLC: has left the MLC towards the next module) and its bank
// (0 field opening, 1 bank A, 2 bank B, -1 not yet located).
// Geometry of the reference: geom.cu:310-790 (howfar), :1152-1286 (hownear),
// init.cu:345-611 (polygons), :1594-1662 (leaf ends per field).
#pragma once
// Included by transport.hpp after Env, Part and the phantom geometry.

namespace gegs {

constexpr int MLC_LEAVES = GEGS_MLC_LEAVES;
constexpr int OUTMLC = -99;
constexpr int PASSMLC = -98;
constexpr uint32_t REG_AIR = 1;          // default_reg
constexpr uint32_t REG_MLC = 3;          // MLC_region

struct P2 {
    float x, y;
};

__device__ __forceinline__ P2 p2(float x, float y) {
    P2 q;
    q.x = x;
    q.y = y;
    return q;
}

__device__ __forceinline__ P2 vtx(const gegs_float2 *poly, int i) { return p2(poly[i].x, poly[i].y); }

// crossing-number test (geom.cu:310-327): a horizontal ray from p to -x
__device__ __forceinline__ bool point_in_polygon(P2 p, const gegs_float2 *poly, int n) {
    bool odd = false;
    int j = n - 1;
    for (int i = 0; i < n; ++i) {
        const P2 a = vtx(poly, i), b = vtx(poly, j);
        if (((a.y < p.y && b.y >= p.y) || (b.y < p.y && a.y >= p.y)) && (a.x <= p.x || b.x <= p.x)) {
            if (a.x + (p.y - a.y) / (b.y - a.y) * (b.x - a.x) < p.x) odd = !odd;
        }
        j = i;
    }
    return odd;
}

// bank of a point in the y-z plane of leaf pair `en` (geom.cu:330-344)
__device__ __forceinline__ int point_in_leaf(P2 p, const gegs_float2 *en) {
    if (point_in_polygon(p, en, 17)) return 1;
    if (point_in_polygon(p, en + 17, 17)) return 2;
    return 0;
}

// leaf whose cross-section holds p (x across the leaves, y = z); -(i+1) in
// the gap after leaf i, OUTMLC outside the leaf rows (geom.cu:346-410)
__device__ __forceinline__ int index_of_leaf(const gegs_mlc &m, P2 p) {
    const float x = p.x / p.y * m.zmin;
    const float z = p.y - m.zmin;
    if (z < 0.0f || z > m.zthick) return OUTMLC;
    if (x < m.x_start || x > m.x_end) return OUTMLC;
    int i = OUTMLC;
    if (x < m.x_start + m.wt) {
        i = 0;
    } else if (x < m.x_mid_1 + m.wt) {
        i = (int)((x - m.x_start - m.wt) / (m.wl_full + m.air_gap));
    } else if (x < m.x_mid_2 + m.wt) {
        const float pair = m.wl_target + m.wl_isocenter + m.air_gap * 2.0f;
        i = (int)((x - m.x_mid_1 - m.wt) / pair);
        if (x - m.x_mid_1 - m.wt - pair * (float)i < m.wl_target + m.air_gap) i = 10 + i * 2;
        else i = 11 + i * 2;
    } else if (x < m.x_end - m.wl_full) {
        i = (int)((x - m.x_mid_2 - m.wt) / (m.wl_full + m.air_gap)) + 50;
    } else if (x <= m.x_end) {
        i = 58;
    }
    int index = OUTMLC;
    if (i >= 0) {
        if (point_in_polygon(p, m.cross_xz + i * 16, 14)) index = i;
        else if (i < MLC_LEAVES - 1) index = point_in_polygon(p, m.cross_xz + (i + 1) * 16, 14) ? i + 1 : -(i + 1);
    }
    return index;
}

// distance along v (not normalised: the result is in units of |v|) to the
// nearest polygon edge, and whether p0 is inside (odd number of crossings);
// an edge is hit for u in [0, 1) so that a vertex counts once (geom.cu:442-484)
__device__ __forceinline__ float ray_polygon(P2 p0, P2 v, const gegs_float2 *poly, int n, bool &inside) {
    float best = INFI_DIST + 1.0f;
    int hits = 0;
    for (int i = 0; i < n; ++i) {
        const int j = i + 1 < n ? i + 1 : 0;
        const P2 a = vtx(poly, i), b = vtx(poly, j);
        const float den = v.x * (b.y - a.y) - v.y * (b.x - a.x);
        if (den != 0.0f) {
            const float d = ((b.x - a.x) * (p0.y - a.y) - (b.y - a.y) * (p0.x - a.x)) / den;
            const float u = ((a.x - p0.x) * v.y - (a.y - p0.y) * v.x) / den;
            if (d >= 0.0f && u >= 0.0f && u < 1.0f) {
                hits++;
                if (d < best) best = d;
            }
        }
    }
    inside = (hits & 1) != 0;
    return best;
}

__device__ __forceinline__ float seg_dist(P2 p, P2 v, P2 w) {
    const float l2 = (v.x - w.x) * (v.x - w.x) + (v.y - w.y) * (v.y - w.y);
    if (l2 == 0.0f) return sqrtf((p.x - v.x) * (p.x - v.x) + (p.y - v.y) * (p.y - v.y));
    const float t = ((p.x - v.x) * (w.x - v.x) + (p.y - v.y) * (w.y - v.y)) / l2;
    if (t < 0.0f) return sqrtf((p.x - v.x) * (p.x - v.x) + (p.y - v.y) * (p.y - v.y));
    if (t > 1.0f) return sqrtf((p.x - w.x) * (p.x - w.x) + (p.y - w.y) * (p.y - w.y));
    const float ax = v.x + t * (w.x - v.x), ay = v.y + t * (w.y - v.y);
    return sqrtf((p.x - ax) * (p.x - ax) + (p.y - ay) * (p.y - ay));
}

// perpendicular distance to the nearest edge (geom.cu:1152-1191)
__device__ __forceinline__ float poly_dist(P2 p, const gegs_float2 *poly, int n) {
    float best = 1e9f;
    for (int i = 0; i < n; ++i) {
        const int j = i + 1 < n ? i + 1 : 0;
        const float d = seg_dist(p, vtx(poly, i), vtx(poly, j));
        if (d < best) best = d;
    }
    return best;
}

// the y-z polygon a particle in leaf pair `en` with bank `bank` is bounded by
__device__ __forceinline__ const gegs_float2 *yz_polygon(const gegs_float2 *en, int bank, int &n) {
    if (bank == 0) { n = 30; return en + 2; }
    n = 17;
    return bank == 2 ? en + 17 : en;
}

// region of a point in the MLC: W inside a leaf, air otherwise (geom.cu:486-516)
__device__ __forceinline__ uint32_t inwhere_mlc(const gegs_mlc &m, const gegs_float2 *endnodes, float qx, float qy,
                                                float qz, int &index, int &bank) {
    uint32_t inew = REG_AIR;
    if (index == OUTMLC) index = index_of_leaf(m, p2(qx, qz));
    if (index >= 0) {
        bank = point_in_leaf(p2(qy, qz), endnodes + index * 34);
        inew = bank > 0 ? REG_MLC : REG_AIR;
    }
    return inew;
}

// distance to the next MLC boundary along the flight and the region entered;
// updates the particle's leaf index, bank and, on leaving, its module
// (geom.cu:518-790)
__device__ __forceinline__ uint32_t howfar_mlc(const Env &env, Part &p, float &t) {
    const gegs_mlc &m = *env.mlc;
    const gegs_float2 *endnodes = env.endnodes + (size_t)p.job * MLC_LEAVES * 34;
    uint32_t inew = p.region;
    int bank = p.bank;
    int index = p.leaf;
    int lim1 = 0, lim2 = MLC_LEAVES;
    if (p.y < -m.rmax || p.y > m.rmax) return 0u;
    const P2 p1 = p2(p.x, p.z), pyz = p2(p.y, p.z);
    const P2 d1 = p2(p.u, p.w), d2 = p2(p.v, p.w);
    bool in_xz, in_yz;
    float dis_xz = ray_polygon(p1, d1, m.boundary_xz, 4, in_xz);
    float dis_yz = ray_polygon(pyz, d2, m.boundary_yz, 4, in_yz);
    if (dis_xz >= INFI_DIST || dis_yz >= INFI_DIST) return 0u;
    const float bound = dis_xz < dis_yz ? dis_xz : dis_yz;
    bool new_region = false;
    if (in_xz && in_yz) {
        // limit the leaf search to the leaves between here and the exit point
        const P2 p3 = p2(p1.x + dis_xz * d1.x, p1.y + (dis_xz - 0.001f) * d1.y);
        lim1 = index_of_leaf(m, p3);
        if (lim1 < 0) lim1 = -lim1;
    } else {
        // outside the MLC: step to its boundary
        inew = 0u;
        if (bound >= INFI_DIST) return inew;
        if (t > bound) {
            t = bound + EPSTEP;
            new_region = true;
        } else {
            return inew;
        }
    }
    if (!new_region) {
        index = p.leaf > -MLC_LEAVES ? p.leaf : index_of_leaf(m, p1);
        if (index >= 0) {
            dis_xz = ray_polygon(p1, d1, m.cross_xz + index * 16, 14, in_xz);
            int nyz;
            const gegs_float2 *yz = yz_polygon(endnodes + index * 34, p.bank, nyz);
            dis_yz = ray_polygon(pyz, d2, yz, nyz, in_yz);
            if (dis_yz > dis_xz) {
                // the side of the leaf comes first: the particle goes into the gap
                if (d1.x < 0.0f && index > 0) index = -index;
                else if (index < MLC_LEAVES - 1) index = -(index + 1);
            }
        } else if (index > -MLC_LEAVES) {
            // in a gap: search the leaves the ray can reach
            int sx = 1, sl = 16;
            if (d1.x < 0.0f) {
                sx = -1;
                sl = -16;
                if (lim1 > 0) lim1 -= 1;
                lim2 = -index + 1;
            } else {
                lim2 = lim1;
                lim1 = -index - 1;
            }
            int i = -index - sx;
            if (lim1 < 0) lim1 = 0;
            if (lim2 > MLC_LEAVES - 1) lim2 = MLC_LEAVES - 1;
            if (i < 0) i = 0;
            if (i > MLC_LEAVES - 1) i = MLC_LEAVES - 1;
            const gegs_float2 *leaf = m.cross_xz + i * 16;
            float tmp;
            do {
                tmp = ray_polygon(p1, d1, leaf, 14, in_xz);
                i += sx;
                leaf += sl;
            } while (i >= lim1 && i <= lim2 && tmp >= INFI_DIST);
            if (tmp < dis_xz) {
                dis_xz = tmp;
                index = i - sx;
            } else {
                index = PASSMLC;
                p.module = env.mlc_next;
            }
        }
        float dmin = dis_xz < dis_yz ? dis_xz : dis_yz;
        bool stepping_out = false;
        if (dmin + EPSTEP >= bound) {
            dmin = bound;
            stepping_out = true;
        }
        if (t > dmin) {
            t = dmin + EPSTEP;
            if (stepping_out) {
                p.leaf = PASSMLC;
                p.module = env.mlc_next;
                return REG_AIR;
            }
        } else {
            return inew;
        }
    }
    // the region at the end of the step
    const float qx = p.x + t * p.u, qy = p.y + t * p.v, qz = p.z + t * p.w;
    inew = inwhere_mlc(m, endnodes, qx, qy, qz, index, bank);
    p.leaf = index;
    p.bank = bank;
    return inew;
}

// distance to the nearest MLC boundary (geom.cu:1193-1286)
__device__ __forceinline__ float hownear_mlc(const Env &env, Part &p) {
    const gegs_mlc &m = *env.mlc;
    const gegs_float2 *endnodes = env.endnodes + (size_t)p.job * MLC_LEAVES * 34;
    const P2 p1 = p2(p.x, p.z), pyz = p2(p.y, p.z);
    float tp_xz = INFI_DIST + 1.0f, tp_yz = INFI_DIST + 1.0f;
    bool in_xz = true, in_yz = true;
    if (p.leaf < -MLC_LEAVES) {
        in_xz = point_in_polygon(p1, m.boundary_xz, 4);
        in_yz = point_in_polygon(pyz, m.boundary_yz, 4);
    }
    const float bxz = poly_dist(p1, m.boundary_xz, 4);
    const float byz = poly_dist(pyz, m.boundary_yz, 4);
    const float bound = bxz < byz ? bxz : byz;
    if (!(in_xz && in_yz)) return bound;
    int index;
    if (p.leaf > PASSMLC) {
        index = p.leaf;
    } else {
        index = index_of_leaf(m, p1);
        p.leaf = index;
    }
    if (index >= 0) {
        tp_xz = poly_dist(p1, m.cross_xz + index * 16, 14);
        if (p.bank < 0) p.bank = point_in_leaf(pyz, endnodes + index * 34);
        int nyz;
        const gegs_float2 *yz = yz_polygon(endnodes + index * 34, p.bank, nyz);
        tp_yz = poly_dist(pyz, yz, nyz);
    } else if (index > -MLC_LEAVES) {
        const int i = -index;
        const float a = poly_dist(p1, m.cross_xz + i * 16, 14);
        tp_xz = poly_dist(p1, m.cross_xz + (i - 1) * 16, 14);
        if (a < tp_xz) tp_xz = a;
    }
    return tp_xz < tp_yz ? tp_xz : tp_yz;
}

}  // namespace gegs
