// blockwedge.hpp -- block and wedge head modules (two BEAMnrc PYRAMIDS).
//
// Each module is the intersection of two prisms of the head frame given by
// their cross-sections: the block by the quadrilaterals blockXZ and blockYZ,
// the wedge by a polygon of 3..6 vertices (goem_Wed) and the rectangle
// boundWed.  The reference builds with ROTATE_Block_XY and ROTATE_Wedge_XY
// (CUDA_EGS.h:28-29): the first polygon is tested with the point (y, z) and
// direction (v, w), the second with (x, z) and (u, w); kept.
// Geometry of the reference: geom.cu:1037-1149 (howfar), :1376-1412 (hownear).
// Device layout of env.bw (18 vertices): [0,4) blockXZ, [4,8) blockYZ,
// [8,14) goem_Wed, [14,18) boundWed.
#pragma once
// Included by transport.hpp after mlc.hpp.

namespace gegs {

constexpr uint32_t REG_WEDGE = 5;         // wedge_reg
constexpr uint32_t REG_BLOCK = 4;         // block_reg
constexpr int BW_VERTS = 18;

// howfar_block / howfar_wedge: a particle outside the module's prisms steps
// onto the nearer one and is taken to be inside; inside both, it leaves at the
// nearer exit for the next module; a ray that misses either prism hands the
// particle to the next module at once.
__device__ __forceinline__ uint32_t howfar_prisms(Part &p, float &t, const gegs_float2 *pa, int na,
                                                  const gegs_float2 *pb, uint32_t reg, int32_t next) {
    bool in_a, in_b;
    const float da = ray_polygon(p2(p.y, p.z), p2(p.v, p.w), pa, na, in_a);
    const float db = ray_polygon(p2(p.x, p.z), p2(p.u, p.w), pb, 4, in_b);
    if (da >= INFI_DIST || db >= INFI_DIST) {
        p.module = next;
        return REG_AIR;
    }
    const float bound = da < db ? da : db;
    if (in_a && in_b) {
        if (t > bound) {
            t = bound + EPSTEP;
            p.module = next;
            return REG_AIR;
        }
        return reg;
    }
    if (t > bound) {
        t = bound + EPSTEP;
        return reg;
    }
    return REG_AIR;
}

__device__ __forceinline__ float hownear_prisms(const Part &p, const gegs_float2 *pa, int na, const gegs_float2 *pb) {
    const float a = poly_dist(p2(p.y, p.z), pa, na);
    const float b = poly_dist(p2(p.x, p.z), pb, 4);
    return a < b ? a : b;
}

// one inlined copy serves both modules: the block (module GEGS_M_BLOCK) and
// the wedge differ only in their vertices, region and next module
__device__ __forceinline__ uint32_t howfar_block_wedge(const Env &env, Part &p, float &t) {
    const bool w = p.module == GEGS_M_WEDGE;
    return howfar_prisms(p, t, env.bw + (w ? 8 : 0), w ? env.wedge_nver : 4, env.bw + (w ? 14 : 4),
                         w ? REG_WEDGE : REG_BLOCK, w ? (int32_t)GEGS_M_PHANTOM : env.block_next);
}

__device__ __forceinline__ float hownear_block_wedge(const Env &env, const Part &p) {
    const bool w = p.module == GEGS_M_WEDGE;
    return hownear_prisms(p, env.bw + (w ? 8 : 0), w ? env.wedge_nver : 4, env.bw + (w ? 14 : 4));
}

}  // namespace gegs
