// jaws.hpp -- secondary collimator jaws (SecJaws, two focused tungsten pairs).
//
// Each pair is a slab of the head frame; the problem is two-dimensional in
// the plane across the jaws.  Per field, 32 vertices (host side, setup of
// the jaw polygons of init.cu:1504-1592):
//   [0,4)   boundMinY: the field rectangle scaled to z = zminy, +-0.5 cm
//   [4,8)   boundMinX: the same at z = zminx
//   [8,12)  boundMaxY: the Y slab (|r| <= rmax, zminy <= z <= zmaxy)
//   [12,16) boundMaxX: the X slab
//   [16,24) jawsGeomY: the -side jaw [0,4), the opening [2,6), the +side jaw [4,8)
//   [24,32) jawsGeomX: the same for X
// The reference builds with ROTATE_SecJaws_XY (CUDA_EGS.h:25): the plane of
// the jaw polygons is (x, z) for the Y pair and (y, z) for the X pair, and
// the entry test against boundMin uses the point (y, x).  Both are kept.
// Geometry of the reference: geom.cu:793-1035 (howfar), :1288-1374 (hownear).
#pragma once
// Included by transport.hpp after mlc.hpp.

namespace gegs {

constexpr uint32_t REG_JAWS = 2;          // secjaws_reg
constexpr int JAW_VERTS = 32;

__device__ __forceinline__ const gegs_float2 *jaw_polys(const Env &env, const Part &p) {
    return env.jaws + (size_t)p.job * JAW_VERTS;
}

// howfar_SecJawsY (Y = true) and howfar_SecJawsX (geom.cu:793-1035)
template <bool Y>
__device__ __forceinline__ uint32_t howfar_jaws(const Env &env, Part &p, float &t) {
    const gegs_float2 *jp = jaw_polys(env, p);
    const gegs_float2 *bmin = jp + (Y ? 0 : 4);
    const gegs_float2 *bmax = jp + (Y ? 8 : 12);
    const gegs_float2 *geom = jp + (Y ? 16 : 24);
    uint32_t inew = 0;
    const P2 p1 = p2(p.y, p.z), q2 = p2(p.x, p.z);
    const P2 d1 = p2(p.v, p.w), e2 = p2(p.u, p.w);
    bool in_a, in_b;
    float dis_xz = ray_polygon(p1, d1, bmax, 4, in_a);
    float dis_yz = ray_polygon(q2, e2, bmax, 4, in_b);
    if (dis_xz >= INFI_DIST || dis_yz >= INFI_DIST) return inew;
    const float bound = dis_xz < dis_yz ? dis_xz : dis_yz;
    // the jaw polygons lie in (x, z) for the Y pair and in (y, z) for the X pair
    const P2 pj = Y ? q2 : p1;
    const P2 dj = Y ? e2 : d1;
    if (in_a && in_b) {
        bool dummy;
        if (point_in_polygon(pj, geom + 2, 4)) {
            // in the opening
            inew = REG_AIR;
            const float da = ray_polygon(pj, dj, geom, 4, dummy);
            const float db = ray_polygon(pj, dj, geom + 4, 4, dummy);
            if (da < INFI_DIST) {
                if (t > da) { inew = REG_JAWS; t = da + EPSTEP; }
            } else if (db < INFI_DIST) {
                if (t > db) { inew = REG_JAWS; t = db + EPSTEP; }
            } else if (t > bound) {
                t = bound + EPSTEP;
                p.module = Y ? GEGS_M_SECJAWX : env.jawx_next;
            }
        } else {
            // in the tungsten
            inew = REG_JAWS;
            const float dopen = ray_polygon(pj, dj, geom + 2, 4, dummy);
            if (dopen < INFI_DIST) {
                // the X pair steps by the boundary distance of the (x, z)
                // plane here (quirk Q8 of geom.cu:992, kept unless
                // GEGS_FIX_Q8 is set: then by the distance to the opening)
                if (t > dopen) {
                    inew = REG_AIR;
                    t = ((Y || (env.fixes & GEGS_FIX_Q8)) ? dopen : dis_yz) + EPSTEP;
                }
            } else if (t > bound) {
                inew = 0;
                t = bound + EPSTEP;
            }
        }
    } else {
        // outside the slab: step onto it and locate the entry point
        inew = REG_AIR;
        if (t > bound) {
            t = bound + EPSTEP;
            P2 q = p2(p.y + t * p.v, p.x + t * p.u);
            if (!point_in_polygon(q, bmin, 4)) {
                inew = 0;
            } else {
                if (Y) q.x = q.y;
                q.y = p.z + t * p.w;
                inew = point_in_polygon(q, geom + 2, 4) ? REG_AIR : REG_JAWS;
            }
        }
    }
    return inew;
}

// hownear_SecJawsY / hownear_SecJawsX (geom.cu:1288-1374)
template <bool Y>
__device__ __forceinline__ float hownear_jaws(const Env &env, const Part &p) {
    const gegs_float2 *jp = jaw_polys(env, p);
    const gegs_float2 *bmax = jp + (Y ? 8 : 12);
    const gegs_float2 *geom = jp + (Y ? 16 : 24);
    const P2 p1 = p2(p.y, p.z), q2 = p2(p.x, p.z);
    const bool in_a = point_in_polygon(p1, bmax, 4);
    const bool in_b = point_in_polygon(q2, bmax, 4);
    const float b1 = poly_dist(p1, bmax, 4);
    const float b2 = poly_dist(q2, bmax, 4);
    if (!(in_a && in_b)) return b1 < b2 ? b1 : b2;
    // the plane of the jaw polygons and the boundary distance of the other one
    const P2 pj = Y ? q2 : p1;
    const float b_own = Y ? b2 : b1, b_other = Y ? b1 : b2;
    float d = poly_dist(pj, geom + 2, 4);
    if (!point_in_polygon(pj, geom + 2, 4) && d > b_own) d = b_own;
    return d < b_other ? d : b_other;
}

}  // namespace gegs
