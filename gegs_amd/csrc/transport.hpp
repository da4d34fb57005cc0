// transport.hpp -- device-side coupled electron/photon transport for gfx950.
//
// Event handlers of the reference's event kernels, one device function per
// event type, shared by the plain per-history kernel (plain.hpp) and the
// event-queued kernels (event.hpp):
//   photon_step   photon.cu:6-292       compton   compton.cu:17-492
//   photo         photo.cu:1-279        rayleigh  rayleigh.cu:4-157
//   pair          pair.cu:14-431        electron  electron.cu:888-2622
//   moller        moller.cu:16-210      brems     brems.cu:36-308
//   bhabha        bhabha.cu:12-172      annih     annih.cu:16-206
//   annih_at_rest annihatrest.cu:10-68  mscat / msdist_pI electron.cu:77-882
//   howfar/hownear_phantom  geom.cu:54-307, geom.cu:1422-1543
// A handler consumes one particle from its event type and hands every
// particle it creates, or forwards to another event type, to a Sink: the
// per-thread stack of the plain kernel or the SoA queues of the event path.
//
// Random numbers: every particle owns a Philox-4x32-10 stream keyed by
// (seed, 64-bit lineage) with a 32-bit block counter; four uniforms per
// block, in (0, 1].  The buffer lives only while the particle is being
// handled; a forwarded particle keeps its lineage and next block, a created
// particle gets a lineage derived from its parent's lineage and stream
// position.  Results therefore do not depend on processing order, which is
// what lets the plain kernel, the event-queued kernels and the adjoint replay
// reproduce each other particle by particle.
//
// Arithmetic is binary32 with explicit float literals except where the
// reference keeps double (prm, the FP64 tallies).  Build with
// -ffp-contract=off and correctly rounded division/sqrt.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Occupancy floor of a kernel: at most 512 / n registers per lane (VGPRs and
// AGPRs together), so that n waves fit on a SIMD (the electron kernel sits
// at the 3-wave limit of 168 VGPRs, the plain kernel at the 2-wave limit).
#ifndef GEGS_WAVES_PER_EU
#define GEGS_WAVES_PER_EU(n) __attribute__((amdgpu_waves_per_eu(n)))
#endif

#include "../../include/gegs.h"
#include "gmath.hpp"

namespace gegs {

constexpr float RM = 0.5110034f;
constexpr double PRM = 0.5110034;
constexpr float RMT2 = 2.0f * RM;
constexpr float RMSQ = RM * RM;
constexpr float VACUUM_STEP = 1e8f;
constexpr double EPSGMFP = 1e-5;
constexpr double EPSEMFP = 1e-5;
constexpr float EPSTEP = 0.0005f;
constexpr float INFI_DIST = 1e5f;
constexpr uint32_t VAC_AND_OTH = 6;
constexpr uint32_t VACUUM = GEGS_VACUUM;
constexpr int MXMED = GEGS_MXMED;
constexpr int MXGE = GEGS_MXGE;
constexpr int MXEKE = GEGS_MXEKE;
constexpr int MXRAYFF = GEGS_MXRAYFF;
constexpr int RAYCDFSIZE = GEGS_RAYCDFSIZE;
constexpr int MAXL_MS = GEGS_MAXL_MS;
constexpr int MAXQ_MS = GEGS_MAXQ_MS;
constexpr int MAXU_MS = GEGS_MAXU_MS;
constexpr float HC_INVERSE = 80.65506856998f;
constexpr float TWICE_HC2 = 0.000307444456f;

// particle_status of the reference (CUDA_EGS.h:416-427)
enum Event : int {
    P_PHOTON_STEP = 0, P_COMPTON = 1, P_PHOTO = 2, P_RAYLEIGH = 3, P_PAIR = 4,
    E_ELECTRON_STEP = 5, E_MOLLER = 6, E_BREMS = 7, E_BHABHA = 8, E_ANNIH = 9,
    N_EVENTS = 10
};

struct Part {
    float e, x, y, z, u, v, w, wt, dnear;
    uint32_t region;
    int32_t charge;
    int32_t status;
    int32_t module;
    uint64_t lin;
    uint32_t blk;
    uint32_t job;        // field of the plan the history belongs to (beam geometry, dose buffer)
    int32_t leaf;        // MLC leaf index (OUTMLC / PASSMLC outside the leaf rows)
    int32_t bank;        // MLC bank: 0 opening, 1 bank A, 2 bank B, -1 not yet located
    int32_t spos;        // bits 0-2: uniforms of block blk-1 already used (4: none
                         // pending); bit 3: a suspended flight that resumes
};

// ------------------------------------------------------------------ Philox
__host__ __device__ __forceinline__ void philox10(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3,
                                         uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t m0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t m1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(m1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(m0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)m1;
        c3 = (uint32_t)m0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__host__ __device__ __forceinline__ uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t primary_lineage(uint64_t hist) {
    return splitmix(hist ^ 0x5851F42D4C957F2Dull);
}

struct Stream {
    uint64_t lin;
    uint32_t blk;            // block after the one being consumed
    int pos;                 // uniforms of the current block already used
    int nchild;
    float b0, b1, b2, b3;    // current block
    float n0, n1, n2, n3;    // block blk, generated ahead of use when `ahead`
    bool ahead;
    uint32_t k0, k1;

    __device__ __forceinline__ void gen(uint32_t block, float &o0, float &o1, float &o2, float &o3) const {
        uint32_t c0 = block, c1 = (uint32_t)lin, c2 = (uint32_t)(lin >> 32), c3 = 0x47454753u;
        philox10(c0, c1, c2, c3, k0, k1);
        const float s = 5.96046448e-08f;
        o0 = (float)((c0 >> 8) + 1u) * s;
        o1 = (float)((c1 >> 8) + 1u) * s;
        o2 = (float)((c2 >> 8) + 1u) * s;
        o3 = (float)((c3 >> 8) + 1u) * s;
    }
    __device__ __forceinline__ void refill() {
        if (ahead) {
            b0 = n0; b1 = n1; b2 = n2; b3 = n3;
            ahead = false;
        } else {
            gen(blk, b0, b1, b2, b3);
        }
        blk += 1;
        pos = 0;
    }
    // Generates the next block before it is needed.  Called by all lanes of
    // a wavefront together once one of them has used up its current block,
    // so that one full-width Philox evaluation serves most lanes, instead of
    // the divergent per-lane refills inside next() running on nearly every
    // iteration with a few lanes each.  The uniforms drawn are unchanged.
    __device__ __forceinline__ void lookahead() {
        if (!ahead) {
            gen(blk, n0, n1, n2, n3);
            ahead = true;
        }
    }
    __device__ __forceinline__ bool starved() const { return pos == 4 && !ahead; }
    // a particle suspended in the middle of a block (spos < 4) resumes at the
    // same uniform: the block is generated again
    __device__ __forceinline__ void begin(const Part &p, uint64_t seed) {
        lin = p.lin;
        blk = p.blk;
        pos = 4;
        nchild = 0;
        ahead = false;
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        if ((p.spos & 7) < 4) {
            blk -= 1;
            refill();
            pos = p.spos & 7;
        }
    }
    __device__ __forceinline__ float next() {
        if (pos == 4) refill();
        const int k = pos++;
        return k == 0 ? b0 : (k == 1 ? b1 : (k == 2 ? b2 : b3));
    }
    __device__ __forceinline__ uint64_t child() {
        const uint64_t at = (uint64_t)blk * 4u + (uint64_t)pos;
        const uint64_t k = (uint64_t)(nchild++);
        return lin ^ splitmix((at << 8) | k);
    }
};

// all lanes of the wavefront must reach this together
__device__ __forceinline__ void wave_lookahead(Stream &s, bool have) {
    if (__any(have && s.starved())) {
        if (have) s.lookahead();
    }
}

// For steps that draw several uniforms (a condensed-history step draws five
// on average): every lane without a block generated ahead makes it as soon as
// any lane lacks one, so a step starts with its current block's rest plus a
// whole block in registers and seldom has to generate one alone mid-step.
// Each block is still generated exactly once; only the lanes that share the
// evaluation change.  All lanes of the wavefront must reach this together.
__device__ __forceinline__ void wave_lookahead_eager(Stream &s, bool have) {
    if (__any(have && !s.ahead)) {
        if (have) s.lookahead();
    }
}

__device__ __forceinline__ uint64_t child_lineage(Stream &s) { return splitmix(s.child()); }

// ------------------------------------------------------------------ environment
// Region records are stored as a class index per region (2 bytes) into a
// table of the distinct records: a CT phantom has a few thousand distinct
// (medium, density) records, so the index volume of a 128x128x100 phantom
// (3.3 MB) stays resident in the L2 of every XCD, where 16-byte records
// (26 MB) would miss to HBM at every voxel crossing.  Values are unchanged.
struct Env {
    const gegs_media_tables *mt;
    const uint16_t *rcls;            // class of region r, 6 + nx*ny*nz entries
    const gegs_region *rtab;         // distinct region records
    const float *xb, *yb, *zb;   // N+2 entries each
    const gegs_float2 *pge, *pgmfp, *pcohe;   // photon rows read at every crossing (mt->ge, gmfp, cohe)
    uint32_t nx, ny, nz;
    uint32_t nclass, nmed;
    const gegs_field *fields;        // beam geometry per job
    uint32_t njob;
    const gegs_mlc *mlc;             // MLC geometry (null: no MLC)
    const gegs_float2 *endnodes;     // leaf ends per job: [job][60][34]
    int32_t mlc_next;                // module after the MLC
    int32_t jawx_next;               // module after the X jaws (the MLC when loaded)
    const gegs_float2 *jaws;         // jaw polygons per job: [job][32]
    const gegs_float2 *bw = nullptr; // block and wedge cross-sections (18 vertices; null: neither loaded)
    int32_t block_next = GEGS_M_PHANTOM;  // module after the block
    int32_t wedge_nver = 0;          // vertices of the wedge cross-section
    uint32_t fixes = 0;              // GEGS_FIX_* corrections of reference quirks (0: as the reference)
    uint64_t seed;
};

// LDS staging of what a voxel crossing reads: the voxel bounds and, when the
// phantom has at most LDS_MAX_CLASSES distinct region records, the record
// table and the photon rows ge/gmfp/cohe of every medium.  Layout:
// [records nclass x 16 B][ge nmed x 8 B][gmfp nmed x MXGE x 8 B][cohe ...]
// [fields njob x 32 B][bounds].
// All threads of the block call stage_env() before any of them tracks a
// particle; the env it returns points at LDS.
constexpr uint32_t LDS_MAX_CLASSES = 2048;

__host__ __device__ inline size_t env_lds_bytes(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nclass, uint32_t nmed,
                                                uint32_t njob, bool tab) {
    size_t b = (size_t)njob * sizeof(gegs_field) + ((size_t)nx + ny + nz + 6) * sizeof(float);
    if (tab) b += (size_t)nclass * 16 + (size_t)nmed * 8 + 2 * (size_t)nmed * MXGE * 8;
    return b;
}

template <bool TAB>
__device__ __forceinline__ void stage_env(Env &env, float *lds) {
    float *p = lds;
    const gegs_region *rtab = env.rtab;
    const gegs_float2 *pge = env.pge, *pgmfp = env.pgmfp, *pcohe = env.pcohe;
    if (TAB) {
        float4 *rt = reinterpret_cast<float4 *>(p);
        const float4 *src = reinterpret_cast<const float4 *>(env.rtab);
        for (uint32_t i = threadIdx.x; i < env.nclass; i += blockDim.x) rt[i] = src[i];
        rtab = reinterpret_cast<const gegs_region *>(rt);
        p += (size_t)env.nclass * 4;
        gegs_float2 *g = reinterpret_cast<gegs_float2 *>(p);
        const uint32_t nrow = env.nmed * MXGE;
        for (uint32_t i = threadIdx.x; i < env.nmed; i += blockDim.x) g[i] = env.mt->ge[i];
        for (uint32_t i = threadIdx.x; i < nrow; i += blockDim.x) {
            g[env.nmed + i] = env.mt->gmfp[i];
            g[env.nmed + nrow + i] = env.mt->cohe[i];
        }
        pge = g;
        pgmfp = g + env.nmed;
        pcohe = g + env.nmed + nrow;
        p += 2 * ((size_t)env.nmed + 2 * nrow);
    }
    {
        const uint32_t nw = env.njob * (uint32_t)(sizeof(gegs_field) / 4);
        const uint32_t *src = reinterpret_cast<const uint32_t *>(env.fields);
        uint32_t *dst = reinterpret_cast<uint32_t *>(p);
        for (uint32_t i = threadIdx.x; i < nw; i += blockDim.x) dst[i] = src[i];
        env.fields = reinterpret_cast<const gegs_field *>(p);
        p += nw;
    }
    const uint32_t nxb = env.nx + 2, nyb = env.ny + 2, nzb = env.nz + 2;
    for (uint32_t i = threadIdx.x; i < nxb; i += blockDim.x) p[i] = env.xb[i];
    for (uint32_t i = threadIdx.x; i < nyb; i += blockDim.x) p[nxb + i] = env.yb[i];
    for (uint32_t i = threadIdx.x; i < nzb; i += blockDim.x) p[nxb + nyb + i] = env.zb[i];
    __syncthreads();
    env.xb = p;
    env.yb = p + nxb;
    env.zb = p + nxb + nyb;
    env.rtab = rtab;
    env.pge = pge;
    env.pgmfp = pgmfp;
    env.pcohe = pcohe;
}

__device__ __forceinline__ gegs_region region_rec(const Env &env, uint32_t r) { return env.rtab[env.rcls[r]]; }

__device__ __forceinline__ const gegs_field &field_of(const Env &env, const Part &p) { return env.fields[p.job]; }

// Scoring policy.  Forward (ADJ = false): every deposit E in voxel v is added
// to the dose buffer in FP64 (escore, scoring.cu:14-24).  Adjoint replay
// (ADJ = true): the deposit is weighted by dL/dD of that voxel and summed per
// particle (acc); the sum goes to the gradient of the history's beamlet.
template <bool ADJ>
struct TallyT {
    double *dose;          // per voxel, FP64 (forward)
    const double *adj;     // dL/dD per voxel (adjoint)
    double acc;            // sum of dL/dD * E over the particle's deposits (adjoint)
    double hist_edep;
    uint32_t steps;
    uint32_t photon_steps, electron_steps, gathers, deposits;

    __device__ __forceinline__ void score(uint32_t region, float energy) {
        if (region < VAC_AND_OTH) return;
        const double v = (double)energy;
        if (ADJ) acc += adj[region - VAC_AND_OTH] * v;
        else if (dose) unsafeAtomicAdd(dose + (region - VAC_AND_OTH), v);
        hist_edep += v;
        deposits++;
    }
};
using Tally = TallyT<false>;
using AdjTally = TallyT<true>;

template <class Sink>
__device__ __forceinline__ void forward(Sink &sink, Stream &s, Part &p, int status) {
    p.status = status;
    p.lin = s.lin;
    p.blk = s.blk;
    p.spos = 4;
    sink.push(p);
}

template <class Sink>
__device__ __forceinline__ void create(Sink &sink, Stream &s, Part &p, int status) {
    p.status = status;
    p.lin = child_lineage(s);
    p.blk = 0;
    p.spos = 4;
    sink.push(p);
}

// ------------------------------------------------------------------ geometry
struct Local {
    float x, y, z, u, v, w;
};

// the transform in binary64 (GEGS_FIX_FP64_FRAME): the field's binary32
// cos/sin pair normalised to unit length, products and sums in binary64,
// one rounding per coordinate
__device__ __forceinline__ void frame64_rotation(const gegs_field &f, double &c, double &s) {
    c = (double)f.cos_angle;
    s = (double)f.sin_angle;
    const double r = sqrt(c * c + s * s);
    c = c / r;
    s = s / r;
}

__device__ __forceinline__ Local phantom_frame64(const gegs_field &f, const Part &p) {
    double x, y, z, u, v, w, c, s;
    frame64_rotation(f, c, s);
    if (f.change_xyz) {
        x = p.y; y = (double)p.z - (double)f.source_to_isocenter; z = p.x;
        u = p.v; v = p.w; w = p.u;
    } else {
        x = p.x; y = p.y; z = (double)p.z - (double)f.source_to_isocenter;
        u = p.u; v = p.v; w = p.w;
    }
    Local q;
    q.x = (float)(x * c - y * s + (double)f.iso_x);
    q.y = (float)(x * s + y * c + (double)f.iso_y);
    q.z = (float)(z + (double)f.iso_z);
    q.u = (float)(u * c - v * s);
    q.v = (float)(u * s + v * c);
    q.w = (float)w;
    return q;
}

// Options of the run that change the flight loops, beside the public
// GEGS_FIX_* bits of Env::fixes: the electron-step algorithms the loaded
// tables select (the reference's compile-time ESA_PII and BCA_EXACT,
// CUDA_EGS.h:35-38), set by the host from the tables.
constexpr uint32_t ENV_PRESTA2 = 1u << 16;     // msdist_pII instead of msdist_pI
constexpr uint32_t ENV_EXACT_BCA = 1u << 17;   // exact boundary crossing (single scattering)
constexpr uint32_t ENV_GENERAL = GEGS_FIX_FP64_FRAME | ENV_PRESTA2 | ENV_EXACT_BCA;

// The flight and plain kernels are instantiated twice: GEN = false for the
// reference's defaults (binary32 frame, PRESTA-I, inexact BCA) with these
// bits known to be zero, so that kernel carries none of their code, and
// GEN = true, which tests them at run time.
template <bool GEN>
__device__ __forceinline__ uint32_t kernel_fixes(uint32_t fixes) {
    return GEN ? fixes : (fixes & ~ENV_GENERAL);
}

__device__ __forceinline__ Local phantom_frame(const gegs_field &f, const Part &p, uint32_t fixes) {
    if (fixes & GEGS_FIX_FP64_FRAME) return phantom_frame64(f, p);
    float x, y, z, u, v, w;
    if (f.change_xyz) {
        x = p.y; y = p.z - f.source_to_isocenter; z = p.x;
        u = p.v; v = p.w; w = p.u;
    } else {
        x = p.x; y = p.y; z = p.z - f.source_to_isocenter;
        u = p.u; v = p.v; w = p.w;
    }
    Local q;
    const float rx = x * f.cos_angle - y * f.sin_angle;
    const float ry = x * f.sin_angle + y * f.cos_angle;
    const float ru = u * f.cos_angle - v * f.sin_angle;
    const float rv = u * f.sin_angle + v * f.cos_angle;
    q.x = rx + f.iso_x;
    q.y = ry + f.iso_y;
    q.z = z + f.iso_z;
    q.u = ru;
    q.v = rv;
    q.w = w;
    return q;
}

__device__ __forceinline__ int locate(float p, uint32_t n, const float *b) {
    if (p < b[0] || p > b[n]) return -1;
    if (n == 1) return 0;
    int lo = 0, hi = (int)n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (p <= b[mid]) hi = mid; else lo = mid;
    }
    return hi - 1;
}

// distance to the next voxel face along the flight direction; returns the
// region entered (0 = outside) and shortens t (geom.cu:54-307)
__device__ __forceinline__ uint32_t howfar_phantom(const Env &env, const Part &p, float &t) {
    const Local q = phantom_frame(field_of(env, p), p, env.fixes);
    const uint32_t nx = env.nx, ny = env.ny, nz = env.nz, nxy = nx * ny;
    const float *xb = env.xb, *yb = env.yb, *zb = env.zb;
    if (p.region >= VAC_AND_OTH) {
        uint32_t ir = p.region - VAC_AND_OTH;
        int iz = (int)(ir / nxy);
        ir -= (uint32_t)iz * nxy;
        int iy = (int)(ir / nx);
        int ix = (int)(ir - (uint32_t)iy * nx);
        // one division per axis, the face ahead chosen by the sign of the
        // direction cosine (the same operands as the reference's two branches)
        uint32_t inew = p.region;
        if (q.u != 0.0f) {
            const bool up = q.u > 0.0f;
            const float d = (xb[ix + (up ? 1 : 0)] - q.x) / q.u;
            if (d <= t) { t = d; inew = up ? ((ix + 1 < (int)nx) ? p.region + 1 : 0u) : ((ix - 1 >= 0) ? p.region - 1 : 0u); }
        }
        if (q.v != 0.0f) {
            const bool up = q.v > 0.0f;
            const float d = (yb[iy + (up ? 1 : 0)] - q.y) / q.v;
            if (d <= t) { t = d; inew = up ? ((iy + 1 < (int)ny) ? p.region + nx : 0u) : ((iy - 1 >= 0) ? p.region - nx : 0u); }
        }
        if (q.w != 0.0f) {
            const bool up = q.w > 0.0f;
            const float d = (zb[iz + (up ? 1 : 0)] - q.z) / q.w;
            if (d <= t) { t = d; inew = up ? ((iz + 1 < (int)nz) ? p.region + nxy : 0u) : ((iz - 1 >= 0) ? p.region - nxy : 0u); }
        }
        if (t < 0.0f) t = EPSTEP;
        return inew;
    }
    // from outside: try the x, y, z entry faces in turn
    float t1 = 0.0f;
    int ix = -1, iy, iz;
    if (q.x <= xb[0] && q.u > 0.0f) { t1 = (xb[0] - q.x) / q.u; ix = 0; }
    else if (q.x >= xb[nx] && q.u < 0.0f) { t1 = (xb[nx] - q.x) / q.u; ix = (int)nx - 1; }
    if (ix >= 0 && t1 <= t) {
        iy = locate(q.y + q.v * t1, ny, yb);
        if (iy >= 0) {
            iz = locate(q.z + q.w * t1, nz, zb);
            if (iz >= 0) { t = t1; return (uint32_t)(ix + iy * (int)nx + iz * (int)nxy) + VAC_AND_OTH; }
        }
    }
    iy = -1;
    if (q.y <= yb[0] && q.v > 0.0f) { t1 = (yb[0] - q.y) / q.v; iy = 0; }
    else if (q.y >= yb[ny] && q.v < 0.0f) { t1 = (yb[ny] - q.y) / q.v; iy = (int)ny - 1; }
    if (iy >= 0 && t1 <= t) {
        ix = locate(q.x + q.u * t1, nx, xb);
        if (ix >= 0) {
            iz = locate(q.z + q.w * t1, nz, zb);
            if (iz >= 0) { t = t1; return (uint32_t)(ix + iy * (int)nx + iz * (int)nxy) + VAC_AND_OTH; }
        }
    }
    iz = -1;
    if (q.z <= zb[0] && q.w > 0.0f) { t1 = (zb[0] - q.z) / q.w; iz = 0; }
    else if (q.z >= zb[nz] && q.w < 0.0f) { t1 = (zb[nz] - q.z) / q.w; iz = (int)nz - 1; }
    if (iz >= 0 && t1 <= t) {
        ix = locate(q.x + q.u * t1, nx, xb);
        if (ix >= 0) {
            iy = locate(q.y + q.v * t1, ny, yb);
            if (iy >= 0) { t = t1; return (uint32_t)(ix + iy * (int)nx + iz * (int)nxy) + VAC_AND_OTH; }
        }
    }
    return 0u;
}

// perpendicular distance to the closest face; keeps the reference's
// region-5 voxel offset (geom.cu:1487-1490), reading the neighbouring voxel
__device__ __forceinline__ float hownear_phantom(const Env &env, const Part &p) {
    const Local q = phantom_frame(field_of(env, p), p, env.fixes);
    const uint32_t nx = env.nx, ny = env.ny, nz = env.nz;
    const float *xb = env.xb, *yb = env.yb, *zb = env.zb;
    // the reference takes the voxel of region r as r - 5 (quirk Q1, geom.cu:1487-1490): the
    // neighbouring voxel; GEGS_FIX_Q1 uses the voxel itself, r - 6, as howfar does
    const uint32_t first = (env.fixes & GEGS_FIX_Q1) ? 6u : 5u;
    if (p.region >= first) {
        uint32_t ir = p.region - first;
        const int iz = (int)(ir / (nx * ny));
        ir -= (uint32_t)iz * nx * ny;
        const int iy = (int)(ir / nx);
        const int ix = (int)(ir - (uint32_t)iy * nx);
        float a, dx, dy, dz;
        a = xb[ix + 1] - q.x; dx = q.x - xb[ix]; if (a < dx) dx = a;
        a = yb[iy + 1] - q.y; dy = q.y - yb[iy]; if (a < dy) dy = a;
        a = zb[iz + 1] - q.z; dz = q.z - zb[iz]; if (a < dz) dz = a;
        float tp = dx < dy ? dx : dy;
        if (tp > dz) tp = dz;
        return tp;
    }
    const float lx = xb[0], ly = yb[0], lz = zb[0];
    float hx = xb[nx], hy = yb[ny], hz = zb[nz];
    if (q.x - lx > 0 && q.x - hx > 0) {} else if (q.x - lx < 0 && q.x - hx < 0) hx = lx; else hx = q.x;
    if (q.y - ly > 0 && q.y - hy > 0) {} else if (q.y - ly < 0 && q.y - hy < 0) hy = ly; else hy = q.y;
    if (q.z - lz > 0 && q.z - hz > 0) {} else if (q.z - lz < 0 && q.z - hz < 0) hz = lz; else hz = q.z;
    const float ddx = q.x - hx, ddy = q.y - hy, ddz = q.z - hz;
    return sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
}

}  // namespace gegs

#include "mlc.hpp"
#include "jaws.hpp"
#include "blockwedge.hpp"

namespace gegs {

// module dispatch of the geometry queries (photon.cu:142-158, electron.cu:1491-1505, :1836-1850)
__device__ __forceinline__ uint32_t howfar_module(const Env &env, Part &p, float &t) {
    if (p.module == GEGS_M_PHANTOM) return howfar_phantom(env, p, t);
    if (p.module == GEGS_M_MLC) return howfar_mlc(env, p, t);
    if (p.module == GEGS_M_SECJAWY) return howfar_jaws<true>(env, p, t);
    if (p.module == GEGS_M_BLOCK || p.module == GEGS_M_WEDGE) return howfar_block_wedge(env, p, t);
    return howfar_jaws<false>(env, p, t);
}

__device__ __forceinline__ float hownear_module(const Env &env, Part &p) {
    if (p.module == GEGS_M_PHANTOM) return hownear_phantom(env, p);
    if (p.module == GEGS_M_MLC) return hownear_mlc(env, p);
    if (p.module == GEGS_M_SECJAWY) return hownear_jaws<true>(env, p);
    if (p.module == GEGS_M_BLOCK || p.module == GEGS_M_WEDGE) return hownear_block_wedge(env, p);
    return hownear_jaws<false>(env, p);
}

// ------------------------------------------------------------------ rotations (utility.cu:94-218)
__device__ __forceinline__ void turn(float costhe, float sinthe, float cosphi, float sinphi, Part &p) {
    const float A = p.u, B = p.v, C = p.w;
    const float s2 = A * A + B * B;
    if (s2 < 1e-20f) {
        p.u = sinthe * cosphi;
        p.v = sinthe * sinphi;
        p.w = C * costhe;
        return;
    }
    const float sp = sqrtf(s2);
    const float us = sinthe * cosphi, vs = sinthe * sinphi;
    const float sd = B / sp, cd = A / sp;
    p.u = C * cd * us - sd * vs + A * costhe;
    p.v = C * sd * us + cd * vs + B * costhe;
    p.w = -sp * us + C * costhe;
}

__device__ __forceinline__ void uphi21(Stream &s, float costhe, float sinthe, float &cosphi, float &sinphi, Part &p) {
    const float r = s.next();
    gsincos2pi(r, sinphi, cosphi);
    turn(costhe, sinthe, cosphi, sinphi, p);
}

__device__ __forceinline__ void azimuth_box(Stream &s, float &cphi, float &sphi) {
    float xp, yp, r2;
    do {
        xp = s.next();
        xp = 2.0f * xp - 1.0f;
        yp = s.next();
        r2 = xp * xp + yp * yp;
    } while (r2 > 1.0f);
    r2 = 1.0f / r2;
    cphi = (xp * xp - yp * yp) * r2;
    sphi = 2.0f * xp * yp * r2;
}

// the reference's rest mass prm is a float symbol (commonBlocks dev.h:220): float arithmetic
__device__ __forceinline__ float to_kinetic(float etot) { return etot - RM; }

// ------------------------------------------------------------------ annihilation at rest
template <class Sink, class T>
__device__ __forceinline__ void annih_at_rest(const Env &env, Stream &s, Part &p, const gegs_region &rd, Sink &sink, T &t) {
    const int nsplit = env.mt->nbr_split;
    float edep = 0.0f;
    if (nsplit > 1) p.wt = p.wt / (float)nsplit;
    p.e = RM;
    p.charge = 0;
    for (int i = 1; i <= nsplit; ++i) {
        float costhe = s.next();
        costhe = 2.0f * costhe - 1.0f;
        const float sinthe = sqrtf(fmaxf(0.0f, (1.0f - costhe) * (1.0f + costhe)));
        float cphi, sphi;
        azimuth_box(s, cphi, sphi);
        p.u = sinthe * cphi;
        p.v = sinthe * sphi;
        p.w = costhe;
        if (p.e < rd.pcut) edep += p.e; else create(sink, s, p, P_PHOTON_STEP);
        p.u = -p.u;
        p.v = -p.v;
        p.w = -p.w;
        if (p.e < rd.pcut) edep += p.e; else create(sink, s, p, P_PHOTON_STEP);
    }
    t.score(p.region, edep * p.wt);
}

// ------------------------------------------------------------------ photon flight (photon.cu:6-292)
// One call of photon_walk_iter() is one photon step of the reference: a new
// number of mean free paths is drawn at every voxel crossing (photon.cu:102-103),
// the step is cut at the next voxel face (geom.cu:143-215), and when the
// remaining path falls below EPSGMFP in the same medium the interaction type
// is chosen (photon.cu:229-271).  The quantities that do not change during a
// flight are held in registers instead of being recomputed or reloaded at
// every crossing:
//   ln E and the medium's interpolated gmfp/cohe rows (reloaded only when the
//   medium changes), the direction in the phantom frame, the voxel indices
//   (stepped by one on a crossing instead of decoded by integer division),
//   and the region record of the current voxel (one gather per crossing).
struct PhotonWalk {
    gegs_region rd;          // region record of p.region
    float gle;               // ln E
    int32_t cmed;            // medium of the cached rows, -1: none
    int32_t lgle;
    float gmfpr0, cohf;      // interpolated gmfp and cohe of medium cmed
    float lu, lv, lw;        // direction in the phantom frame
    int32_t ix, iy, iz;      // voxel indices, valid when p.region >= VAC_AND_OTH
    float cohfac;            // Rayleigh factor of the last step (read by photon_interaction)
};

// photon_walk_iter's outcomes: the flight ended (absorbed below the cut,
// discarded, or handed on), goes on, or reached its interaction point
constexpr int PW_END = 0;
constexpr int PW_ON = 1;
constexpr int PW_INTERACT = 2;

__device__ __forceinline__ void local_position(const gegs_field &f, const Part &p, uint32_t fixes, float &qx, float &qy,
                                               float &qz) {
    // the position part of phantom_frame(), operation for operation
    if (fixes & GEGS_FIX_FP64_FRAME) {
        const Local q = phantom_frame64(f, p);
        qx = q.x;
        qy = q.y;
        qz = q.z;
        return;
    }
    float x, y, z;
    if (f.change_xyz) {
        x = p.y; y = p.z - f.source_to_isocenter; z = p.x;
    } else {
        x = p.x; y = p.y; z = p.z - f.source_to_isocenter;
    }
    const float rx = x * f.cos_angle - y * f.sin_angle;
    const float ry = x * f.sin_angle + y * f.cos_angle;
    qx = rx + f.iso_x;
    qy = ry + f.iso_y;
    qz = z + f.iso_z;
}

__device__ __forceinline__ void decode_voxel(const Env &env, uint32_t region, int32_t &ix, int32_t &iy, int32_t &iz) {
    uint32_t ir = region - VAC_AND_OTH;
    const uint32_t nxy = env.nx * env.ny;
    iz = (int32_t)(ir / nxy);
    ir -= (uint32_t)iz * nxy;
    iy = (int32_t)(ir / env.nx);
    ix = (int32_t)(ir - (uint32_t)iy * env.nx);
}

__device__ __forceinline__ void photon_walk_begin(const Env &env, const Part &p, PhotonWalk &w) {
    w.rd = region_rec(env, p.region);
    w.gle = glog(p.e);
    w.cmed = -1;
    w.lgle = 0;
    w.gmfpr0 = w.cohf = 0.0f;
    const Local q = phantom_frame(field_of(env, p), p, env.fixes);
    w.lu = q.u;
    w.lv = q.v;
    w.lw = q.w;
    w.ix = w.iy = w.iz = 0;
    if (p.region >= VAC_AND_OTH) decode_voxel(env, p.region, w.ix, w.iy, w.iz);
}

// One free-flight segment: PW_ON when the photon crossed a boundary and flies
// on, PW_INTERACT at the interaction point (photon_interaction then chooses
// the interaction from the same stream; the event path defers that choice
// until the wavefront hands its ended flights on together), PW_END otherwise.
template <class Sink, class T>
__device__ __forceinline__ int photon_walk_iter(const Env &env, Stream &s, Part &p, PhotonWalk &w, Sink &sink, T &t) {
    const gegs_media_tables *mt = env.mt;
    t.photon_steps++;
    t.steps++;
    gegs_region rd = w.rd;
    if (p.e <= rd.pcut) { t.score(p.region, p.e * p.wt); return PW_END; }
    if (p.wt <= 0.0f) return PW_END;
    // Inside the phantom the face the ray leaves the voxel through does not
    // depend on the sampled path, so it is found first and the class of the
    // voxel behind it is requested before the uniform is drawn; the load is
    // in flight while the step length is computed.  howfar_phantom() takes
    // the faces in x, y, z order with `d <= t`: the voxel is left through the
    // last face of minimal distance when that distance is <= the sampled
    // step, which is what the comparison below does.
    const bool inside = p.region >= VAC_AND_OTH;
    float dmin = 0.0f;
    uint32_t ncand = 0;
    int axis = -1;
    uint16_t cls_next = 0;
    if (inside) {
        float qx, qy, qz;
        local_position(field_of(env, p), p, env.fixes, qx, qy, qz);
        // One division per axis: the face ahead is bound index + 1 for a
        // positive direction cosine and + 0 for a negative one, the same
        // operands the reference divides in either case, so lanes moving
        // either way share one instruction stream instead of two branches.
        const uint32_t nx = env.nx, nxy = env.nx * env.ny;
        if (w.lu != 0.0f) {
            const bool up = w.lu > 0.0f;
            dmin = (env.xb[w.ix + (up ? 1 : 0)] - qx) / w.lu;
            ncand = up ? ((w.ix + 1 < (int)nx) ? p.region + 1 : 0u) : ((w.ix - 1 >= 0) ? p.region - 1 : 0u);
            axis = up ? 0 : 1;
        }
        if (w.lv != 0.0f) {
            const bool up = w.lv > 0.0f;
            const float d = (env.yb[w.iy + (up ? 1 : 0)] - qy) / w.lv;
            if (axis < 0 || d <= dmin) {
                dmin = d;
                ncand = up ? ((w.iy + 1 < (int)env.ny) ? p.region + nx : 0u) : ((w.iy - 1 >= 0) ? p.region - nx : 0u);
                axis = up ? 2 : 3;
            }
        }
        if (w.lw != 0.0f) {
            const bool up = w.lw > 0.0f;
            const float d = (env.zb[w.iz + (up ? 1 : 0)] - qz) / w.lw;
            if (axis < 0 || d <= dmin) {
                dmin = d;
                ncand = up ? ((w.iz + 1 < (int)env.nz) ? p.region + nxy : 0u) : ((w.iz - 1 >= 0) ? p.region - nxy : 0u);
                axis = up ? 4 : 5;
            }
        }
        if (axis >= 0) cls_next = env.rcls[ncand];
    }
    float r1 = s.next();
    float dpmfp = -glog(r1);
    const float gle = w.gle;
    int lgle = 0;
    float tstep, gmfp_val = 0.0f, cohfac = 0.0f;
    if (rd.med == VACUUM) {
        tstep = VACUUM_STEP;
    } else {
        if ((int32_t)rd.med != w.cmed) {
            const gegs_float2 ge = env.pge[rd.med];
            w.lgle = (int)(ge.x + ge.y * gle);
            const gegs_float2 gm = env.pgmfp[rd.med * MXGE + w.lgle];
            const gegs_float2 ch = env.pcohe[rd.med * MXGE + w.lgle];
            w.gmfpr0 = gm.x + gm.y * gle;
            w.cohf = ch.x + ch.y * gle;
            w.cmed = rd.med;
        }
        lgle = w.lgle;
        gmfp_val = w.gmfpr0 / rd.rhof;
        if (rd.flags & GEGS_F_RAYLEIGH) {
            cohfac = w.cohf;
            gmfp_val *= cohfac;
        }
        tstep = gmfp_val * dpmfp;
    }
    uint32_t nreg;
    bool crossed = false;
    if (inside) {
        nreg = p.region;
        if (axis >= 0 && dmin <= tstep) {
            tstep = dmin;
            nreg = ncand;
            crossed = true;
        }
        if (tstep < 0.0f) tstep = EPSTEP;
    } else {
        nreg = howfar_module(env, p, tstep);
    }
    int idisc = 0;
    if (nreg == 0) idisc = (rd.med == VACUUM || p.leaf == PASSMLC) ? 1 : -1;
    if (idisc > 0) { p.region = 0; return PW_END; }
    p.x += p.u * tstep;
    p.y += p.v * tstep;
    p.z += p.w * tstep;
    if (rd.med != VACUUM) dpmfp = fmaxf(0.0f, dpmfp - tstep / gmfp_val);
    const uint32_t old_med = rd.med;
    if (nreg != p.region) {
        if (crossed) {
            w.ix += axis == 0 ? 1 : (axis == 1 ? -1 : 0);
            w.iy += axis == 2 ? 1 : (axis == 3 ? -1 : 0);
            w.iz += axis == 4 ? 1 : (axis == 5 ? -1 : 0);
            rd = env.rtab[cls_next];
        } else {
            if (nreg >= VAC_AND_OTH) decode_voxel(env, nreg, w.ix, w.iy, w.iz);
            rd = region_rec(env, nreg);
        }
        p.region = nreg;
        w.rd = rd;
        t.gathers++;
    }
    if (p.e <= rd.pcut || idisc < 0) { t.score(p.region, p.e * p.wt); return PW_END; }
    if (rd.med != old_med || rd.med == VACUUM || (double)dpmfp >= EPSGMFP) return PW_ON;
    // the interaction point: the medium is the one the step started in, so
    // w.lgle and w.gle index its rows; the Rayleigh factor is the step's
    w.cohfac = cohfac;
    return PW_INTERACT;
}

// The interaction at the end of a flight (photon.cu:229-271): Rayleigh with
// probability 1 - cohfac in a region with Rayleigh on, else pair, Compton or
// photo-absorption from the branching ratios at ln E.  Called right after
// photon_walk_iter returned PW_INTERACT, or later with the lane's stream,
// particle and walk state untouched in between: the uniforms are the same.
template <class Sink>
__device__ __forceinline__ void photon_interaction(const Env &env, Stream &s, Part &p, const PhotonWalk &w, Sink &sink) {
    const gegs_media_tables *mt = env.mt;
    const gegs_region &rd = w.rd;
    float r1;
    if (rd.flags & GEGS_F_RAYLEIGH) {
        r1 = s.next();
        if (r1 < 1.0f - w.cohfac) { forward(sink, s, p, P_RAYLEIGH); return; }
    }
    const gegs_float2 g1 = mt->gbr1[rd.med * MXGE + w.lgle];
    const gegs_float2 g2 = mt->gbr2[rd.med * MXGE + w.lgle];
    const float gbr1 = g1.x + g1.y * w.gle;
    const float gbr2 = g2.x + g2.y * w.gle;
    r1 = s.next();
    if (r1 <= gbr1 && p.e > 2.0f * RM) forward(sink, s, p, P_PAIR);
    else if (r1 < gbr2) forward(sink, s, p, P_COMPTON);
    else forward(sink, s, p, P_PHOTO);
}

// a flight interrupted between two crossings: the particle keeps its stream
// position so that it resumes with the next uniform
template <class Sink>
__device__ __forceinline__ void suspend(Sink &sink, Stream &s, Part &p, int status) {
    p.status = status;
    p.lin = s.lin;
    p.blk = s.blk;
    p.spos = s.pos | 8;
    sink.push(p);
}

// VOTE: the lanes that step together generate their look-ahead Philox blocks
// together (a vote among them); false where lanes step different particles
// unsynchronised and no vote can be shared.  The uniforms are the same.
template <bool VOTE = true, class Sink, class T>
__device__ __forceinline__ void photon_step(const Env &env, Stream &s, Part &p, Sink &sink, T &t) {
    PhotonWalk w;
    photon_walk_begin(env, p, w);
    for (;;) {
        if (VOTE) wave_lookahead(s, true);
        const int r = photon_walk_iter(env, s, p, w, sink, t);
        if (r == PW_ON) continue;
        if (r == PW_INTERACT) photon_interaction(env, s, p, w, sink);
        break;
    }
}

// ------------------------------------------------------------------ atomic relaxation (relax.cu:14-238)
constexpr int MXSHELL = GEGS_MXSHELL;
constexpr int MXTRANS = GEGS_MXTRANS;
constexpr int MXEDGE = GEGS_MXEDGE;
constexpr int MXMDSH = GEGS_MXMDSH;
constexpr int MXVAC = GEGS_MXVAC;
constexpr int N_EII_BINS = GEGS_N_EII_BINS;
constexpr double RELAX_CUTOFF = 0.001;

// transitions of each shell (commonBlocks dev.h:260-262): [first, last]
// indices into the element's probabilities, and the final state of each:
// < 10 fluorescence into that shell, 1x Coster-Kronig into shell x, ab Auger
// leaving vacancies in a and b
__device__ __forceinline__ int first_transition(int shell) {
    return shell == 1 ? 1 : (shell == 2 ? 20 : (shell == 3 ? 27 : (shell == 4 ? 33 : 38)));
}
__device__ __forceinline__ int last_transition(int shell) {
    return shell == 1 ? 19 : (shell == 2 ? 26 : (shell == 3 ? 32 : (shell == 4 ? 37 : 39)));
}
__device__ __forceinline__ int final_state(int k) {
    // two entries per 32-bit word of a wavefront-uniform table
    constexpr uint32_t FS[20] = {
        4u | 3u << 16, 5u | 6u << 16, 202u | 302u << 16, 402u | 404u << 16, 403u | 303u << 16,
        502u | 503u << 16, 504u | 602u << 16, 603u | 604u << 16, 505u | 605u << 16, 606u | 13u << 16,
        14u | 5u << 16, 6u | 505u << 16, 605u | 606u << 16, 14u | 5u << 16, 6u | 505u << 16,
        605u | 606u << 16, 5u | 6u << 16, 505u | 605u << 16, 606u | 6u << 16, 606u};
    const uint32_t w = FS[(k - 1) >> 1];
    return (int)(((k - 1) & 1) ? (w >> 16) : (w & 0xffffu));
}

// isotropic direction of a relaxation particle (relax.cu:62-90)
__device__ __forceinline__ void relax_direction(Stream &s, Part &p) {
    float eta = s.next();
    eta = 2 * eta - 1;
    p.w = eta;
    eta = (1 - eta) * (1 + eta);
    if (eta > 1e-20f) {
        eta = sqrtf(eta);
        float xphi, yphi, rhophi2;
        do {
            xphi = s.next();
            xphi = 2 * xphi - 1;
            yphi = s.next();
            rhophi2 = xphi * xphi + yphi * yphi;
        } while (rhophi2 > 1);
        rhophi2 = 1 / rhophi2;
        p.u = eta * (xphi * xphi - yphi * yphi) * rhophi2;
        p.v = eta * 2 * xphi * yphi * rhophi2;
    } else {
        p.u = 0.0f;
        p.v = 0.0f;
        p.w = 1.0f;
    }
}

// fill a vacancy in shell n (1 K .. 6 N) of element iz carrying `energy`:
// fluorescence photons and Auger / Coster-Kronig electrons above the cuts are
// created, the rest goes to edep; a cascade deeper than MXVAC vacancies (an
// out-of-bounds write in the reference) deposits the vacancy locally
template <class Sink>
__device__ __forceinline__ void relax(const Env &env, Stream &s, float energy, int n, int iz, float &edep, Part p,
                                      const gegs_region &rd, Sink &sink) {
    const gegs_media_tables *mt = env.mt;
    if (n < 1 || n > MXSHELL) return;
    float e_cut = fminf(rd.ecut - RM, rd.pcut);
    e_cut = ((double)e_cut > RELAX_CUTOFF) ? e_cut : (float)RELAX_CUTOFF;
    if (energy <= e_cut) {
        edep += energy;
        return;
    }
    // vacancy stack, shells and energies, in registers: 4 bits per shell
    uint32_t vac = (uint32_t)n;
    float ea[MXVAC];
    int nvac = 1;
    ea[0] = energy;
    const float *be = mt->binding_energies + MXSHELL * (iz - 1);
    for (;;) {
        const int shell = (int)((vac >> (4 * (nvac - 1))) & 15u);
        const float ei = ea[nvac - 1];
        nvac -= 1;
        if (ei <= e_cut) {
            edep += ei;
            if (nvac > 0) continue;
            break;
        }
        if (shell == MXSHELL) {
            p.e = ei + RM;
            p.charge = -1;
            relax_direction(s, p);
            if (p.e < rd.ecut) edep += p.e - RM;
            else create(sink, s, p, E_ELECTRON_STEP);
            if (nvac > 0) continue;
            break;
        }
        float eta = s.next();
        int k;
        const int klast = last_transition(shell) - 1;
        for (k = first_transition(shell); k <= klast; k++) {
            eta = eta - mt->relaxation_prob[MXTRANS * (iz - 1) + k - 1];
            if (eta <= 0.0f) break;
        }
        int fin = final_state(k);
        int iql;
        float elcut, ex;
        if (nvac + (fin < 100 ? 1 : 2) > MXVAC) {
            edep += ei;
            if (nvac > 0) continue;
            break;
        }
        if (fin < 100) {
            if (fin < 10) {
                iql = 0;
                elcut = rd.pcut;
            } else {
                fin = fin - 10;
                iql = -1;
                elcut = rd.ecut - RM;
            }
            const float ef = be[fin - 1];
            ex = ei - ef;
            nvac += 1;
            vac = (vac & ~(15u << (4 * (nvac - 1)))) | ((uint32_t)fin << (4 * (nvac - 1)));
            ea[nvac - 1] = ef;
        } else {
            const int f1 = fin / 100;
            const int f2 = fin - f1 * 100;
            nvac += 1;
            vac = (vac & ~(15u << (4 * (nvac - 1)))) | ((uint32_t)f1 << (4 * (nvac - 1)));
            ea[nvac - 1] = be[f1 - 1];
            nvac += 1;
            vac = (vac & ~(15u << (4 * (nvac - 1)))) | ((uint32_t)f2 << (4 * (nvac - 1)));
            ea[nvac - 1] = be[f2 - 1];
            iql = -1;
            ex = ei - ea[nvac - 1] - ea[nvac - 2];
            elcut = rd.ecut - RM;
        }
        if (ex <= elcut) {
            edep += ex;
        } else {
            p.charge = iql;
            p.e = iql == 0 ? ex : ex + RM;
            relax_direction(s, p);
            if (iql == 0) {
                if (p.e < rd.pcut) edep += p.e;
                else create(sink, s, p, P_PHOTON_STEP);
            } else {
                if (p.e < rd.ecut) edep += p.e - RM;
                else create(sink, s, p, E_ELECTRON_STEP);
            }
        }
    }
}

// ------------------------------------------------------------------ Compton (compton.cu:17-492)
// ibcmpl from the region flags (compton.cu:87-95)
__device__ __forceinline__ int compton_mode(uint32_t flags) {
    const uint32_t f = flags & 0x000Eu;
    return f == 0x0002u ? 1 : (f == 0x0006u ? 2 : (f == 0x000Au ? 3 : 0));
}

template <class Sink, class T>
__device__ __forceinline__ void compton(const Env &env, Stream &s, Part &p, Sink &sink, T &t) {
    const gegs_media_tables *mt = env.mt;
    const gegs_region rd = region_rec(env, p.region);
    float edep = 0.0f;
    Part e1 = p;
    const int medium = rd.med;
    const float peig = p.e;
    const float ko = peig / RM;
    const float broi = 1 + 2 * ko;
    const int ibcmpl = compton_mode(rd.flags);
    bool first = true;
    // declared inside the resample loop in the reference and read again on a
    // resample without being set: kept from the previous pass
    float bro = 0.0f, br = 0.0f, temp = 0.0f, sinthe = 0.0f, costhe = 0.0f;
    float alph1 = 0.0f, alpha = 0.0f, rejmax = 0.0f;
    for (int loops = 0;; loops++) {
        if (loops > 100) break;
        float uj = 0.0f, jo = 0.0f;
        int j = 0;
        if (ibcmpl > 0) {
            float rnno17 = s.next();
            rnno17 = 1 + rnno17 * mt->n_shell[medium];
            int i = (int)rnno17;
            if (rnno17 > mt->eno_array[medium * MXMDSH + i - 1]) i = mt->eno_atbin_array[medium * MXMDSH + i - 1];
            j = mt->shell_array[medium * MXMDSH + i - 1];
            uj = mt->be_array[j - 1];
            if (ko <= uj) {
                if (ibcmpl == 1) { forward(sink, s, p, P_PHOTON_STEP); break; }
                continue;
            }
            jo = mt->jo_array[j - 1];
        }
        // the pass counter br_loops is never incremented in the reference
        // (quirk Q3, compton.cu:134-141); GEGS_FIX_Q3 applies its cap
        int br_loops = 0;
        do {
            if ((env.fixes & GEGS_FIX_Q3) && br_loops++ > 100) break;
            if (ko > 2) {
                if (first) {
                    alph1 = glog(broi);
                    bro = 1 / broi;
                    alpha = alph1 + ko * (broi + 1) * bro * bro;
                }
                float aux, rejf3, r19;
                int n = 0;
                do {
                    const float r15 = s.next();
                    const float r16 = s.next();
                    if ((r15 * alpha) < alph1) br = gexp(alph1 * r16) * bro;
                    else br = sqrtf(r16 * broi * broi + (1 - r16)) * bro;
                    temp = (1 - br) / (ko * br);
                    sinthe = fmaxf(0.0f, temp * (2 - temp));
                    aux = 1 + br * br;
                    rejf3 = aux - br * sinthe;
                    r19 = s.next();
                    if (++n > 100) break;
                } while (r19 * aux > rejf3);
            } else {
                if (first) {
                    bro = (float)(1.0 / (double)broi);
                    rejmax = broi + bro;
                }
                float r16, rejf3;
                int n = 0;
                do {
                    const float r15 = s.next();
                    r16 = s.next();
                    br = bro + (1 - bro) * r15;
                    temp = (1 - br) / (ko * br);
                    sinthe = fmaxf(0.0f, temp * (2 - temp));
                    rejf3 = 1 + br * br - br * sinthe;
                    if (n++ > 100) break;
                } while (r16 * br * rejmax > rejf3);
            }
            first = false;
        } while (br < bro || br > 1);
        costhe = 1 - temp;
        if (ibcmpl == 0) {
            uj = 0;
        } else {
            // p_z range rejection and sampling (compton.cu:228-405)
            float aux = ko * (ko - uj) * temp;
            float aux1 = 2 * aux + uj * uj;
            float pzmax = aux - uj;
            if ((pzmax < 0) && (pzmax * pzmax >= aux1)) {
                if (ibcmpl == 1) { forward(sink, s, p, P_PHOTON_STEP); break; }
                continue;
            }
            pzmax = pzmax / sqrtf(aux1);
            const float qc2 = 1 + br * br - 2 * br * costhe;
            const float qc = sqrtf(qc2);
            float af, fmax1, fpz;
            bool rejected = false;
            if (pzmax > 1) {
                pzmax = 1;
                af = 0;
                fmax1 = 1;
                fpz = 1;
            } else {
                float aux3 = 1 + 2 * jo * fabsf(pzmax);
                float aux4 = (float)(0.5 * (double)(1 - aux3 * aux3));
                fpz = (float)(0.5 * (double)gexp(aux4));
                af = qc * (1 + br * (br - costhe) / qc2);
                if (af < 0) {
                    if (pzmax > 0) fpz = 1.0f - fpz;
                    const float eta = s.next();
                    if (eta > fpz) rejected = true;
                    af = 0;
                    fmax1 = 1;
                } else {
                    float fpz1;
                    if ((double)pzmax < -0.15) {
                        fmax1 = (float)(1.0 - (double)af * 0.15);
                        fpz1 = fpz * fmax1 * jo;
                    } else if ((double)pzmax < 0.15) {
                        fmax1 = 1 + af * pzmax;
                        aux3 = (float)(1.0 / (1.0 + 0.33267252734 * (double)aux3));
                        aux4 = (float)((double)(fpz * aux3) *
                                       (0.3480242 + (double)aux3 * (-0.0958798 + (double)aux3 * 0.7478556)) +
                                       (double)mt->erfjo_array[j - 1]);
                        if (pzmax > 0) {
                            fpz1 = (float)((double)((1 - fmax1 * fpz) * jo) - 0.62665706866 * (double)af * (double)aux4);
                            fpz = 1 - fpz;
                        } else {
                            fpz1 = (float)((double)(fmax1 * fpz * jo) - 0.62665706866 * (double)af * (double)aux4);
                        }
                    } else {
                        fmax1 = (float)(1.0 + (double)af * 0.15);
                        fpz1 = (1 - fmax1 * fpz) * jo;
                        fpz = 1 - fpz;
                    }
                    const float eta = s.next();
                    if (eta * jo > fpz1) rejected = true;
                }
            }
            if (rejected) {
                if (ibcmpl == 1) { forward(sink, s, p, P_PHOTON_STEP); break; }
                continue;
            }
            float pz = 0.0f;
            for (int n = 0;; n++) {
                if (n > 100) break;
                if (ibcmpl != 2) {
                    float r18 = s.next();
                    r18 = r18 * fpz;
                    if (r18 < 0.5f) {
                        r18 = (double)(2 * r18) > 1e-30 ? 2 * r18 : (float)1e-30;
                        pz = (float)(0.5 * (double)(1 - sqrtf(1 - 2 * glog(r18))) / (double)jo);
                    } else {
                        r18 = 2 * (1 - r18);
                        pz = (float)(0.5 * (double)(sqrtf(1 - 2 * glog(r18)) - 1) / (double)jo);
                    }
                    if (fabsf(pz) > 1) continue;
                    if ((double)pz < 0.15) {
                        const float frej = (double)pz < -0.15 ? (float)(1.0 - (double)af * 0.15) : 1 + af * pz;
                        const float eta = s.next();
                        if (eta * fmax1 > frej) continue;
                    }
                } else {
                    pz = 0;
                    uj = 0;
                }
                break;
            }
            if ((double)fabsf(pz) < 0.01) {
                br = br * (1 + pz * (qc + (br * br - costhe) * pz));
            } else {
                aux = 1 - pz * pz * br * costhe;
                aux1 = 1 - pz * pz * br * br;
                const float aux2 = qc2 - br * br * pz * pz * sinthe;
                if ((double)aux2 > 1e-10) br = br / aux1 * (aux + pz * sqrtf(aux2));
            }
            uj = uj * RM;
        }
        const float pesg = br * peig;
        const float pese = peig - pesg - uj + RM;
        sinthe = sqrtf(sinthe);
        p.e = pesg;
        float sinphi, cosphi;
        uphi21(s, costhe, sinthe, cosphi, sinphi, p);
        if (p.e < rd.pcut) edep += p.e; else create(sink, s, p, P_PHOTON_STEP);
        const float aux = 1 + br * br - 2 * br * costhe;
        if ((double)aux > 1e-8) {
            costhe = (1 - br * costhe) / sqrtf(aux);
            sinthe = (1 - costhe) * (1 + costhe);
            sinthe = sinthe > 0 ? -sqrtf(sinthe) : 0.0f;
        } else {
            costhe = 0.0f;
            sinthe = -1.0f;
        }
        e1.e = pese;
        e1.charge = -1;
        turn(costhe, sinthe, cosphi, sinphi, e1);
        if (e1.e < rd.ecut) edep += e1.e - RM; else create(sink, s, e1, E_ELECTRON_STEP);
        if (ibcmpl == 1 || ibcmpl == 3) {
            if ((double)uj > RELAX_CUTOFF) relax(env, s, uj, mt->shn_array[j - 1], mt->iz_array[j - 1], edep, p, rd, sink);
            else edep += uj;
        }
        break;
    }
    t.score(p.region, edep * p.wt);
}

// ------------------------------------------------------------------ photo-electric (photo.cu:1-279)
__device__ __forceinline__ void sauter_angle(Stream &s, Part &p) {
    // Sauter distribution (photo.cu:197-244)
    const float beta = sqrtf((p.e - RM) * (p.e + RM)) / p.e;
    const float gamma = p.e / RM;
    const float alpha = 0.5f * gamma - 0.5f + 1.0f / gamma;
    const float ratio = beta / alpha;
    float xi, r2, sinth2, costhe;
    do {
        float rn = s.next();
        rn = 2.0f * rn - 1.0f;
        if (ratio <= 0.2f) {
            const float fk = rn + 0.5f * ratio * (1.0f - rn) * (1.0f + rn);
            if (gamma < 100.0f) {
                costhe = (beta + fk) / (1.0f + beta * fk);
            } else if (fk > 0.0f) {
                costhe = 1.0f - (1.0f - fk) * (gamma - 3.0f) /
                         (2.0f * (1.0f + fk) * (gamma - 1.0f) * (gamma - 1.0f) * (gamma - 1.0f));
            } else {
                costhe = (beta + fk) / (1.0f + beta * fk);
            }
            xi = (1.0f + beta * fk) * gamma * gamma;
        } else {
            xi = gamma * gamma * (1.0f + alpha * (sqrtf(1.0f + ratio * (2.0f * rn + ratio)) - 1.0f));
            costhe = (1.0f - 1.0f / xi) / beta;
        }
        sinth2 = fmaxf(0.0f, (1.0f - costhe) * (1.0f + costhe));
        r2 = s.next();
    } while (r2 > 0.5f * (1.0f + gamma) * sinth2 * xi / gamma);
    float sinphi, cosphi;
    uphi21(s, costhe, sqrtf(sinth2), cosphi, sinphi, p);
}

template <class Sink, class T>
__device__ __forceinline__ void photo(const Env &env, Stream &s, Part &p, Sink &sink, T &t) {
    const gegs_media_tables *mt = env.mt;
    const gegs_region rd = region_rec(env, p.region);
    float edep = 0.0f;
    const int medium = rd.med;
    const float peig = p.e;
    if (peig < mt->edge_energies[1]) {   // quirk Q6: the second edge of hydrogen
        p.charge = -1;
        p.e = peig + RM;
        if (p.e < rd.ecut) t.score(p.region, (p.e - RM) * p.wt);
        else create(sink, s, p, E_ELECTRON_STEP);
        return;
    }
    const float gle = glog(p.e);
    bool do_relax = false;
    float e_vac = 0.0f;
    int iz = 0, k = 0;
    if (rd.flags & GEGS_F_ATOMIC_RELAXATION) {
        // the element (photo.cu:87-154), then the shell (photo.cu:160-183)
        const int nne = mt->nne[medium];
        if (nne == 1) {
            iz = (int)((double)mt->zelem[medium] + 0.5);
        } else {
            float probs[GEGS_MXEL];
            float aux = peig * peig;
            const float aux1 = aux * peig;
            aux = aux * sqrtf(peig);
            float sigtot = 0;
            for (k = 0; k < nne; k++) {
                const int z = (int)((double)mt->zelem[MXMED * k + medium] + 0.5);
                const int e0 = MXEDGE * (z - 1);
                float sigma;
                if (peig > mt->edge_energies[e0]) {
                    sigma = (mt->edge_a[e0] + mt->edge_b[e0] / peig + mt->edge_c[e0] / aux + mt->edge_d[e0] / aux1) / peig;
                } else {
                    int j;
                    const int ne = mt->edge_number[z - 1];
                    for (j = 2; j <= ne; j++)
                        if (peig >= mt->edge_energies[e0 + j - 1]) break;
                    sigma = mt->edge_a[e0 + j - 1] +
                            gle * (mt->edge_b[e0 + j - 1] + gle * (mt->edge_c[e0 + j - 1] + gle * mt->edge_d[e0 + j - 1]));
                    sigma = gexp(sigma);
                }
                sigma = sigma * mt->pz[MXMED * k + medium];
                sigtot = sigtot + sigma;
                probs[k] = sigma;
            }
            float br = s.next();
            br = br * sigtot;
            for (k = 0; k < nne; k++) {
                br = br - probs[k];
                if (br <= 0) break;
            }
            // past the last element only by rounding (an out-of-range read in the reference)
            if (k == nne) k = nne - 1;
            iz = (int)((double)mt->zelem[MXMED * k + medium] + 0.5);
        }
        const float *be = mt->binding_energies + MXSHELL * (iz - 1);
        if (peig <= be[5]) {
            edep += peig;
        } else {
            float br = s.next();
            for (k = 1; k <= 5; k++) {
                if (peig > be[k - 1]) {
                    const float tmp = mt->interaction_prob[MXSHELL * (iz - 1) + k - 1];
                    if (br < tmp) break;
                    br = (br - tmp) / (1 - tmp);
                }
            }
            e_vac = be[k - 1];
            p.e = peig - e_vac + RM;
            do_relax = true;
            p.charge = -1;
        }
    } else {
        p.e = peig + RM;
        p.charge = -1;
    }
    if (p.charge == -1) {
        if (p.e > rd.ecut) {
            if (rd.flags & GEGS_F_PHOTOELECTRON_ANGLE) sauter_angle(s, p);
            create(sink, s, p, E_ELECTRON_STEP);
        } else {
            edep += p.e - RM;
        }
    }
    if (do_relax) relax(env, s, e_vac, k, iz, edep, p, rd, sink);
    t.score(p.region, edep * p.wt);
}

// ------------------------------------------------------------------ Rayleigh
template <class Sink, class T>
__device__ __forceinline__ void rayleigh(const Env &env, Stream &s, Part &p, Sink &sink, T &t) {
    const gegs_media_tables *mt = env.mt;
    const gegs_region rd = region_rec(env, p.region);
    const float gle = glog(p.e);
    const gegs_float2 ge = mt->ge[rd.med];
    const int lgle = (int)(ge.x + ge.y * gle);
    const gegs_float2 pm = mt->pmax[rd.med * MXGE + lgle];
    const float pmax_val = pm.x + pm.y * gle;
    const float xmax = HC_INVERSE * p.e;
    const uint32_t *iarr = mt->i_array + rd.med * RAYCDFSIZE;
    const gegs_rayleigh *rdat = mt->rayleigh_data + rd.med * MXRAYFF;
    float costhe, costhe2;
    for (;;) {
        float xv;
        for (;;) {
            const float r1 = s.next();
            float temp = r1 * pmax_val;
            const int ibin = (int)(temp * (float)(RAYCDFSIZE - 1));
            int ib = (int)iarr[ibin] - 1;
            const int next_ib = (int)iarr[ibin + 1] - 1;
            if (next_ib > ib) {
                while (!(temp < rdat[ib + 1].fcum || ib >= RAYCDFSIZE - 2)) ib++;
            }
            const gegs_rayleigh rb = rdat[ib];
            temp = (temp - rb.fcum) * rb.c_array;
            xv = rb.xgrid * gexp(glog(1.0f + temp) * rb.b_array);
            if (xv < xmax) break;
        }
        const float r2 = s.next();
        xv = xv / p.e;
        costhe = 1.0f - TWICE_HC2 * xv * xv;
        costhe2 = costhe * costhe;
        if (2.0f * r2 < 1.0f + costhe2) break;
    }
    float cosphi, sinphi;
    uphi21(s, costhe, sqrtf(1.0f - costhe2), cosphi, sinphi, p);
    if (p.e < rd.pcut) t.score(p.region, p.e * p.wt); else create(sink, s, p, P_PHOTON_STEP);
}

// ------------------------------------------------------------------ pair production
template <class Sink, class T>
__device__ __forceinline__ void pair(const Env &env, Stream &s, Part &p, Sink &sink, T &t) {
    const gegs_media_tables *mt = env.mt;
    const gegs_region rd = region_rec(env, p.region);
    const int m8 = 8 * (int)rd.med;
    float edep = 0.0f;
    Part p1 = p;
    const float eig = p.e;
    float ese1, ese2;
    int q1, q2;
    if (eig <= 2.1f) {
        float r = s.next();
        // prm + 0.5*rnno30*(eig - 2*prm) with the double literal 0.5 (pair.cu:89)
        ese2 = (float)((double)RM + 0.5 * (double)r * (double)(eig - 2.0f * RM));
        r = s.next();
        ese1 = eig - ese2;
        q1 = r < 0.5f ? -1 : 1;
    } else {
        float amax, bmax, aux1;
        int l;
        if (eig < 50.0f) {
            l = 5;
            const float delta = 4.0f * mt->delcm[rd.med] / eig;
            if (delta < 1.0f) {
                amax = mt->dl1[m8 + l - 1] + delta * (mt->dl2[m8 + l - 1] + delta * mt->dl3[m8 + l - 1]);
                bmax = mt->dl1[m8 + l] + delta * (mt->dl2[m8 + l] + delta * mt->dl3[m8 + l]);
            } else {
                const float a2 = glog(delta + mt->dl6[m8 + l - 1]);
                amax = mt->dl4[m8 + l - 1] + mt->dl5[m8 + l - 1] * a2;
                bmax = mt->dl4[m8 + l] + mt->dl5[m8 + l] * a2;
            }
            aux1 = 1.0f - RMT2 / eig;
            aux1 = aux1 * aux1;
            aux1 = aux1 * amax / 3.0f;
            aux1 = aux1 / (bmax + aux1);
        } else {
            l = 7;
            amax = mt->dl1[m8 + l - 1];
            bmax = mt->dl1[m8 + l];
            // bpar(1, medium) read at 8 * medium (quirk Q2, pair.cu:137); GEGS_FIX_Q2 reads 2 * medium
            const int b1 = ((env.fixes & GEGS_FIX_Q2) ? 2 : 8) * (int)rd.med;
            aux1 = mt->bpar[2 * rd.med + 1] * (1.0f - mt->bpar[b1] * RM / eig);
        }
        const float del0 = eig * mt->delcm[rd.med];
        const float eavail = eig - RMT2;
        float r30, rejmax, rejf, eminus;
        do {
            r30 = s.next();
            float r31 = s.next();
            float br;
            int l1;
            if (r30 > aux1) {
                br = 0.5f * r31;
                rejmax = bmax;
                l1 = l + 1;
            } else {
                r30 = s.next();
                r30 = fmaxf(r30, r31);
                r31 = s.next();
                br = 0.5f * (1.0f - fmaxf(r30, r31));
                rejmax = amax;
                l1 = l;
            }
            eminus = br * eavail + RM;
            const float delta = del0 / (eminus * (eig - eminus));
            if (delta < 1.0f)
                rejf = mt->dl1[m8 + l1 - 1] + delta * (mt->dl2[m8 + l1 - 1] + delta * mt->dl3[m8 + l1 - 1]);
            else
                rejf = mt->dl4[m8 + l1 - 1] + mt->dl5[m8 + l1 - 1] * glog(delta + mt->dl6[m8 + l1 - 1]);
            r30 = s.next();
        } while (r30 * rejmax > rejf);
        ese2 = eminus;
        ese1 = eig - ese2;
        q1 = s.next() < 0.5f ? -1 : 1;
    }
    q2 = -q1;
    p.e = ese1;
    p1.e = ese2;
    p.charge = q1;
    p1.charge = q2;
    float cosphi = 0.0f, sinphi = 0.0f;
    // iprdst = 1: leading-order angular distribution for each lepton
    for (int k = 1; k <= 2; ++k) {
        const float ese = k == 1 ? ese1 : ese2;
        const float pse = sqrtf(fmaxf(0.0f, (ese - RM) * (ese + RM)));
        float costhe = s.next();
        costhe = 1.0f - 2.0f * costhe;
        float sinthe = RM * sqrtf((1.0f - costhe) * (1.0f + costhe)) / (pse * costhe + ese);
        costhe = (ese * costhe + pse) / (pse * costhe + ese);
        Part &q = k == 1 ? p : p1;
        if (k == 1) {
            uphi21(s, costhe, sinthe, cosphi, sinphi, q);
        } else {
            turn(costhe, -sinthe, cosphi, sinphi, q);
        }
        if (q.e < rd.ecut) {
            edep += to_kinetic(q.e);
            if (q.charge > 0) annih_at_rest(env, s, q, rd, sink, t);
        } else {
            create(sink, s, q, E_ELECTRON_STEP);
        }
    }
    t.score(p.region, edep * p.wt);
}

// ------------------------------------------------------------------ multiple scattering
__device__ __forceinline__ void mscat(const gegs_media_tables *mt, Stream &s, float lambda, float chia2, float q1,
                             float &cost, float &sint) {
    if (lambda <= 13.8f) {
        const float sprob = s.next();
        const float explambda = gexp(-lambda);
        if (sprob < explambda) { cost = 1.0f; sint = 0.0f; return; }
        float wsum = (1.0f + lambda) * explambda;
        if (sprob < wsum) {
            float xi = s.next();
            xi = 2.0f * chia2 * xi / (1.0f - xi + chia2);
            cost = 1.0f - xi;
            sint = sqrtf(xi * (2.0f - xi));
            return;
        }
        if (lambda <= 1.0f) {
            float wprob = explambda;
            wsum = explambda;
            cost = 1.0f;
            sint = 0.0f;
            int icount = 0;
            do {
                if (++icount > 20) break;
                wprob = wprob * lambda / (float)icount;
                wsum = wsum + wprob;
                float xi = s.next();
                xi = 2.0f * chia2 * xi / (1.0f - xi + chia2);
                const float cosz = 1.0f - xi;
                float sinz = xi * (2.0f - xi);
                if (sinz > 1.0e-20f) {
                    sinz = sqrtf(sinz);
                    xi = s.next();
                    float sphi, cphi;
                    gsincos2pi(xi, sphi, cphi);
                    cost = cost * cosz - sint * sinz * cphi;
                    sint = sqrtf(fmaxf(0.0f, (1.0f - cost) * (1.0f + cost)));
                }
            } while (wsum <= sprob);
            return;
        }
    }
    if (lambda <= 1e5f) {
        const float ll = glog(lambda);
        float ai = ll * mt->dllambi;
        int msi = (int)ai;
        ai = ai - (float)msi;
        float xi = s.next();
        if (xi < ai) msi = msi + 1;
        int msj;
        if (q1 < 1e-3f) {
            msj = 0;
        } else if (q1 < 0.5f) {
            float aj = q1 * mt->dqmsi;
            msj = (int)aj;
            aj = aj - (float)msj;
            xi = s.next();
            if (xi < aj) msj = msj + 1;
        } else {
            msj = MAXQ_MS;
        }
        float omega2;
        if (ll < 2.2299f)
            omega2 = chia2 * (lambda + 4.0f) * (1.347006f + ll * (0.209364f - ll * (0.45525f - ll * (0.50142f - 0.081234f * ll))));
        else
            omega2 = chia2 * (lambda + 4.0f) * (-2.77164f + ll * (2.94874f - ll * (0.1535754f - ll * 0.00552888f)));
        xi = s.next();
        float ak = xi * (float)MAXU_MS;
        int msk = (int)ak;
        ak = ak - (float)msk;
        const int stride_k = (MAXQ_MS + 1) * (MAXL_MS + 1);
        const int off = msj * (MAXL_MS + 1) + msi;
        if (ak > mt->wms[msk * stride_k + off]) msk = mt->ims[msk * stride_k + off];
        const float a = mt->fms[msk * stride_k + off];
        float u = mt->ums[msk * stride_k + off];
        const float du = mt->ums[(msk + 1) * stride_k + off] - u;
        xi = s.next();
        if (fabsf(a) < 0.2f) {
            const float x1 = 0.5f * (1.0f - xi) * a;
            u = u + xi * du * (1.0f + x1 * (1.0f - xi * a));
        } else {
            u = u - du / a * (1.0f - sqrtf(1.0f + xi * a * (2.0f + a)));
        }
        xi = omega2 * u / (1.0f + 0.5f * omega2 - u);
        if (xi > 1.99999f) xi = 1.99999f;
        cost = 1.0f - xi;
        sint = sqrtf(xi * (2.0f - xi));
        return;
    }
    cost = 1.0f;
    sint = 0.0f;
}

// msdist_pII (electron.cu:282-614), spin off: two multiple-scattering
// angles, one per half step, combined with random substep lengths b, c and
// the energy-loss corrections of PRESTA-II (Kawrakow & Bielajew 1998)
__device__ __forceinline__ float msdist_pII(const gegs_media_tables *mt, Stream &s, float eloss, float tustep, Part &p,
                                            const gegs_region &rd) {
    const int m = rd.med;
    const float e0 = p.e - RM;
    const float blc = mt->blcc[m];
    const float xcc = mt->xcc[m];
    float e = e0 - 0.5f * eloss;
    const float tau = e / RM;
    const float tau2 = tau * tau;
    const float epsilon = eloss / e0;
    const float epsilonp = eloss / e;
    // the reference's denominator (24 tau2 + 48 tau + 72), kept
    e = e * (1.0f - epsilonp * epsilonp * (6.0f + 10.0f * tau + 5.0f * tau2) / (24.0f * tau2 + 48.0f * tau + 72.0f));
    const float p2 = e * (e + RMT2);
    const float beta2 = p2 / (p2 + RMSQ);
    const float chia2 = xcc / p2 / (4.0f * blc);
    float lambda = 0.5f * tustep * rd.rhof * blc / beta2;
    float temp = epsilonp / ((tau + 1.0f) * (tau + 2.0f));
    float temp1 = 0.166666f * (4.0f + tau * (6.0f + tau * (7.0f + tau * (4.0f + tau)))) * temp * temp;
    lambda = lambda * (1.0f - temp1);
    float gamma = 1.0f;
    lambda = lambda / (1.0f + chia2);
    const float chilog = glog(1.0f + 1.0f / chia2);
    const float q1 = 2.0f * chia2 * (chilog * (1.0f + chia2) - 1.0f);
    gamma = 6.0f * chia2 * (1.0f + chia2) * (chilog * (1.0f + 2.0f * chia2) - 2.0f) / q1 * gamma;
    float xi = q1 * lambda;
    float w1, sint1, w2, sint2, cphi1, sphi1, cphi2, sphi2;
    mscat(mt, s, lambda, chia2, xi, w1, sint1);
    azimuth_box(s, cphi1, sphi1);
    mscat(mt, s, lambda, chia2, xi, w2, sint2);
    azimuth_box(s, cphi2, sphi2);
    const float u2 = sint2 * cphi2;
    const float v2 = sint2 * sphi2;
    float u2p = w1 * u2 + sint1 * w2;
    float us = u2p * cphi1 - v2 * sphi1;
    float vs = u2p * sphi1 + v2 * cphi1;
    float ws = w1 * w2 - sint1 * u2;
    xi = 2.0f * xi;
    float eta = s.next();
    eta = sqrtf(eta);
    const float eta1 = 0.5f * (1.0f - eta);
    float delta = 0.9082483f - (0.1020621f - 0.0263747f * gamma) * xi;
    temp1 = 2.0f + tau;
    temp = (2.0f + tau * temp1) / ((tau + 1.0f) * temp1);
    temp = temp - (tau + 1.0f) / ((tau + 2.0f) * (chilog * (1.0f + chia2) - 1.0f));
    temp = temp * epsilonp;
    temp1 = 1.0f - temp;
    delta = delta + 0.40824829f * (epsilon * (tau + 1.0f) /
                                   ((tau + 2.0f) * (chilog * (1.0f + chia2) - 1.0f) * (chilog * (1.0f + 2.0f * chia2) - 2.0f)) -
                                   0.25f * temp * temp);
    const float b = eta * delta;
    const float cc = eta * (1.0f - delta);
    const float w1v2 = w1 * v2;
    float ut = b * sint1 * cphi1 + cc * (cphi1 * u2 - sphi1 * w1v2) + eta1 * us * temp1;
    float vt = b * sint1 * sphi1 + cc * (sphi1 * u2 + cphi1 * w1v2) + eta1 * vs * temp1;
    float wt = eta1 * (1.0f + temp) + b * w1 + cc * w2 + eta1 * ws * temp1;
    const float ustep = tustep * sqrtf(ut * ut + vt * vt + wt * wt);
    const float sint02 = p.u * p.u + p.v * p.v;
    if (sint02 > 1e-20f) {
        const float sint0 = sqrtf(sint02);
        const float cphi0 = p.u / sint0;
        const float sphi0 = p.v / sint0;
        u2p = p.w * us + sint0 * ws;
        ws = p.w * ws - sint0 * us;
        us = u2p * cphi0 - vs * sphi0;
        vs = u2p * sphi0 + vs * cphi0;
        u2p = p.w * ut + sint0 * wt;
        wt = p.w * wt - sint0 * ut;
        ut = u2p * cphi0 - vt * sphi0;
        vt = u2p * sphi0 + vt * cphi0;
    } else {
        wt = p.w * wt;
        ws = p.w * ws;
    }
    p.x = p.x + tustep * ut;
    p.y = p.y + tustep * vt;
    p.z = p.z + tustep * wt;
    p.u = us;
    p.v = vs;
    p.w = ws;
    return ustep;
}

// sscat (electron.cu:245-279), spin off: one screened-Rutherford scattering
__device__ __forceinline__ void sscat(Stream &s, float chia2, float &cost, float &sint) {
    float xi = s.next();
    xi = 2.0f * chia2 * xi / (1.0f - xi + chia2);
    cost = 1.0f - xi;
    sint = sqrtf(xi * (2.0f - xi));
}

__device__ __forceinline__ float msdist_pI(const gegs_media_tables *mt, Stream &s, float eloss, float tustep, Part &p,
                                  const gegs_region &rd) {
    const int m = rd.med;
    const float e0 = p.e - RM;
    const float blc = mt->blcc[m];
    const float xcc = mt->xcc[m];
    const float e = e0 - 0.5f * eloss;
    const float p2 = e * (e + RMT2);
    const float beta2 = p2 / (p2 + RMSQ);
    float lambda = tustep * rd.rhof * blc / beta2;
    const float factor = 1.0f / (1.0f + 0.9784671f * e);
    float eps = eloss / e0;
    eps = eps / (1.0f - 0.5f * eps);
    const float temp = 0.25f * (1.0f - factor * (1.0f - 0.333333f * factor)) * eps * eps;
    lambda = lambda * (1.0f + temp);
    const float chia2 = xcc / p2 / (4.0f * blc);
    lambda = lambda / (1.0f + chia2);
    const float chilog = glog(1.0f + 1.0f / chia2);
    const float q1 = 2.0f * chia2 * (chilog * (1.0f + chia2) - 1.0f);
    const float xi = q1 * lambda;
    float ws, sint;
    mscat(mt, s, lambda, chia2, xi, ws, sint);
    float cphi, sphi;
    azimuth_box(s, cphi, sphi);
    float us = sint * cphi, vs = sint * sphi;
    float z;
    if (xi < 0.1f) z = 1.0f - xi * (0.5f - xi * (0.166666667f - 0.041666667f * xi));
    else z = (1.0f - gexp(-xi)) / xi;
    float r = 0.5f * sint;
    float r2 = r * r;
    const float z2 = z * z;
    const float r2max = 1.0f - z2;
    if (r2max < r2) {
        r2 = r2max;
        r = sqrtf(r2);
    }
    float ut = r * cphi, vt = r * sphi, wt = z;
    const float ustep = sqrtf(z2 + r2) * tustep;
    const float s02 = p.u * p.u + p.v * p.v;
    if (s02 > 1e-20f) {
        const float s0 = sqrtf(s02);
        const float cp0 = p.u / s0, sp0 = p.v / s0;
        float u2p = p.w * us + s0 * ws;
        ws = p.w * ws - s0 * us;
        us = u2p * cp0 - vs * sp0;
        vs = u2p * sp0 + vs * cp0;
        u2p = p.w * ut + s0 * wt;
        wt = p.w * wt - s0 * ut;
        ut = u2p * cp0 - vt * sp0;
        vt = u2p * sp0 + vt * cp0;
    } else {
        wt = p.w * wt;
        ws = p.w * ws;
    }
    p.x = p.x + tustep * ut;
    p.y = p.y + tustep * vt;
    p.z = p.z + tustep * wt;
    p.u = us;
    p.v = vs;
    p.w = ws;
    return ustep;
}

// energy lost over a path (the reference's $COMPUTE-ELOSS-G)
__device__ __forceinline__ float eloss_over(const gegs_media_tables *mt, const gegs_float2 *dedx, int qel, int m,
                                   float range, float rhof, float step, float eke, float elke, int lelke) {
    const float *rep = mt->range_ep;
    float tuss = range - rep[(m * MXEKE + lelke - 1) * 2 + qel] / rhof;
    float de;
    if (tuss >= step) {
        const gegs_float2 dd = dedx[lelke - 1];
        const float dmid = dd.x + dd.y * elke;
        const float aux = dd.y / dmid;
        de = dmid * step * rhof;
        const float fe = de / eke;
        de = de * (1.0f - 0.5f * fe * aux * (1.0f - 0.333333f * fe * (aux - 1.0f - 0.25f * fe * (2.0f - aux * (4.0f - aux)))));
        return de;
    }
    tuss = (range - step) * rhof;
    if (tuss <= 0.0f) return eke - mt->te[m] * 0.99f;
    int l = lelke;
    while (tuss < rep[(m * MXEKE + l - 1) * 2 + qel]) l = l - 1;
    const gegs_float2 ek = mt->eke01[m];
    const float elk = ((float)(l + 1) - ek.x) / ek.y;
    const float ekt = mt->e_array[m * MXEKE + l];
    tuss = (rep[(m * MXEKE + l) * 2 + qel] - tuss) / rhof;
    const gegs_float2 dd = dedx[l - 1];
    const float dmid = dd.x + dd.y * elk;
    const float aux = dd.y / dmid;
    de = dmid * tuss * rhof;
    const float fe = de / ekt;
    de = de * (1.0f - 0.5f * fe * aux * (1.0f - 0.333333f * fe * (aux - 1.0f - 0.25f * fe * (2.0f - aux * (4.0f - aux)))));
    return de + eke - ekt;
}

// path length from eke down to e1 within one interpolation bin (the
// reference's $COMPUTE-DRANGE)
__device__ __forceinline__ float path_in_bin(const gegs_float2 *dedx, float eke, float elke, float e1, float elk1, int lbin) {
    const float fe = 1.0f - e1 / eke;
    const float elmid = 0.5f * (elke + elk1 + 0.25f * fe * fe * (1.0f + fe * (1.0f + 0.875f * fe)));
    const gegs_float2 dd = dedx[lbin - 1];
    float dmid = dd.x + dd.y * elmid;
    dmid = 1.0f / dmid;
    float aux = dd.y * dmid;
    const float tq = fe / (2.0f - fe);
    aux = aux * (1.0f + 2.0f * aux) * tq * tq / 6.0f;
    return fe * eke * dmid * (1.0f + aux);
}

// ------------------------------------------------------------------ condensed-history electron step
// state of an electron that persists between condensed-history steps
struct EState {
    float elke, demfp, sig0, total_tstep;
    int lelke;
    bool compute_tstep, is_tstep;
    __device__ __forceinline__ void init() {
        elke = demfp = sig0 = total_tstep = 0.0f;
        lelke = 0;
        compute_tstep = true;
        is_tstep = true;
    }
};

template <class Sink, class T>
__device__ __forceinline__ bool electron_iter(const Env &env, Stream &s, Part &p, EState &st, Sink &sink, T &t) {
    // one condensed-history step; true: the electron continues stepping
    const gegs_media_tables *mt = env.mt;
    float &elke = st.elke, &demfp = st.demfp, &sig0 = st.sig0, &total_tstep = st.total_tstep;
    int &lelke = st.lelke;
    bool &compute_tstep = st.compute_tstep, &is_tstep = st.is_tstep;
    uint32_t irl = p.region;
    gegs_region rd = region_rec(env, irl);
    int medium = rd.med;
    const float rhof = rd.rhof;
    const int lelec = p.charge;
    const int qel = (1 + lelec) / 2;
    float eie = p.e;
    t.electron_steps++;
    t.steps++;
    if (eie <= rd.ecut) {
        t.score(p.region, to_kinetic(p.e) * p.wt);
        if (p.charge > 0) annih_at_rest(env, s, p, rd, sink, t);
        return false;
    }
    if (p.wt <= 0.0f) return false;
    float eke = eie - RM;
    const int mq = medium == (int)VACUUM ? 0 : medium;
    const gegs_float2 *dedx = mt->dedx + qel * MXEKE * MXMED + mq * MXEKE;
    if (is_tstep) {
        compute_tstep = true;
        if (medium != (int)VACUUM) {
            float rn = s.next();
            if (rn == 0.0f) rn = 1.0e-30f;
            demfp = fmaxf(-glog(rn), (float)EPSEMFP);
            elke = glog(eke);
            const gegs_float2 ek = mt->eke01[medium];
            lelke = (int)(ek.x + ek.y * elke);
            if (mt->sig_ismonotone[medium * 2 + qel]) {
                const gegs_float2 sd = mt->sig[qel * MXEKE * MXMED + medium * MXEKE + lelke - 1];
                const gegs_float2 dd = dedx[lelke - 1];
                float sigf = sd.x + sd.y * elke;
                sigf /= dd.x + dd.y * elke;
                sig0 = sigf;
            } else {
                sig0 = mt->sig_e[qel * MXMED + medium];
            }
        }
    }
    float sig_val = 0.0f, dedx_val = 0.0f, tstep = 0.0f, ustep = 0.0f, tustep = 0.0f;
    float tvstep = 0.0f, vstep = 0.0f, range = 0.0f, de = 0.0f;
    float blccl = 0.0f, xccl = 0.0f, beta2 = 0.0f, chia2 = 0.0f, xi = 0.0f, tperp = 0.0f;
    bool callhowfar = false, domultiple = false, dosingle = false, callmsdist = false;
    // electron-step options of the loaded tables (known zero in the
    // reference-default kernels)
    const bool exact_bca = (env.fixes & ENV_EXACT_BCA) != 0;
    if (medium == (int)VACUUM) {
        tstep = VACUUM_STEP;
        ustep = tstep;
        tustep = ustep;
        callhowfar = true;
    } else {
        const gegs_float2 ek = mt->eke01[medium];
        sig_val = sig0;
        if (sig_val <= 0.0f) {
            tstep = VACUUM_STEP;
            sig0 = 1.0e-15f;
        } else {
            if (compute_tstep) {
                const float ekef = eke - demfp / sig_val;
                if (ekef <= mt->e_array[medium * MXEKE]) {
                    tstep = VACUUM_STEP;
                } else {
                    const float elkef = glog(ekef);
                    const int lelkef = (int)(ek.x + ek.y * elkef);
                    if (lelkef == lelke) {
                        tstep = path_in_bin(dedx, eke, elke, ekef, elkef, lelke);
                    } else {
                        const float ekei = mt->e_array[medium * MXEKE + lelke - 1];
                        const float elkei = ((float)lelke - ek.x) / ek.y;
                        const float tuss = path_in_bin(dedx, eke, elke, ekei, elkei, lelke);
                        const float ekej = mt->e_array[medium * MXEKE + lelkef];
                        const float elkej = ((float)(lelkef + 1) - ek.x) / ek.y;
                        tstep = path_in_bin(dedx, ekej, elkej, ekef, elkef, lelkef);
                        tstep = tstep + tuss + mt->range_ep[(medium * MXEKE + lelke - 1) * 2 + qel] -
                                mt->range_ep[(medium * MXEKE + lelkef) * 2 + qel];
                    }
                }
                total_tstep = tstep;
                compute_tstep = false;
            }
            tstep = total_tstep / rhof;
        }
        const gegs_float2 dd = dedx[lelke - 1];
        dedx_val = rhof * (dd.x + dd.y * elke);
        const gegs_float2 tm = mt->tmxs[medium * MXEKE + lelke - 1];
        float tmxs_val = tm.x + tm.y * elke;
        tmxs_val /= rhof;
        const float ekei = mt->e_array[medium * MXEKE + lelke - 1];
        const float elkei = ((float)lelke - ek.x) / ek.y;
        range = path_in_bin(dedx, eke, elke, ekei, elkei, lelke);
        range = (range + mt->range_ep[(medium * MXEKE + lelke - 1) * 2 + qel]) / rhof;
        tmxs_val = fminf(tmxs_val, mt->smaxir[medium]);
        tustep = fminf(tstep, fminf(tmxs_val, range));
        tperp = hownear_module(env, p);
        p.dnear = tperp;
        // range rejection compares the total energy with e_max_rr (quirk Q7, electron.cu:1523);
        // GEGS_FIX_Q7 compares the kinetic energy
        const float e_rr = (env.fixes & GEGS_FIX_Q7) ? eke : p.e;
        if (mt->i_do_rr[medium] == 1 && e_rr < mt->e_max_rr[medium] && tperp >= range) {
            t.score(p.region, to_kinetic(p.e) * p.wt);
            if (lelec > 0) annih_at_rest(env, s, p, rd, sink, t);
            return false;
        }
        blccl = rhof * mt->blcc[medium];
        xccl = rhof * mt->xcc[medium];
        const float p2 = eke * (eke + RMT2);
        beta2 = p2 / (p2 + RMSQ);
        const float ssmfp = beta2 / blccl;
        const float skindepth = mt->skindepth_for_bca * ssmfp;
        tustep = fminf(tustep, fmaxf(tperp, skindepth));
        if (tustep <= tperp && (!exact_bca || tustep > skindepth)) {
            callmsdist = true;
            de = eloss_over(mt, dedx, qel, medium, range, rhof, tustep, eke, elke, lelke);
            tvstep = tustep;
            if (env.fixes & ENV_PRESTA2) ustep = msdist_pII(mt, s, de, tustep, p, rd);
            else ustep = msdist_pI(mt, s, de, tustep, p, rd);
        } else if (exact_bca) {
            // exact BCA (electron.cu:1718-1752): cross the boundary in
            // single-scattering mode; the distance to the next elastic
            // scattering, in mean free paths, shortens the step
            const float rnnoss = s.next();
            const float lambda = rnnoss < 1.0f ? -glog(1.0f - rnnoss) : __builtin_inff();
            const float tmp = eke / RM + 1.0f;
            const float lambda_max = 0.5f * blccl * RM / dedx_val * tmp * tmp * tmp;
            if (lambda >= 0.0f && lambda_max > 0.0f) {
                float tuss;
                if (lambda < lambda_max) tuss = lambda * ssmfp * (1.0f - 0.5f * lambda / lambda_max);
                else tuss = 0.5f * lambda * ssmfp;
                if (tuss < tustep) {
                    tustep = tuss;
                    dosingle = true;
                }
            } else {
                return false;      // the reference drops the electron (electron.cu:1742-1748)
            }
            ustep = tustep;
            callhowfar = !(ustep < tperp);
        } else {
            // inexact boundary-crossing algorithm
            domultiple = true;
            const float ekems = eke - 0.5f * tustep * dedx_val;
            const float p2m = ekems * (ekems + RMT2);
            beta2 = p2m / (p2m + RMSQ);
            chia2 = xccl / (4.0f * blccl * p2m);
            xi = 0.5f * xccl / p2m / beta2 * tustep;
            xi = xi * (glog(1.0f + 1.0f / chia2) - 1.0f / (1.0f + chia2));
            if (xi < 0.1f) ustep = tustep * (1.0f - xi * (0.5f - xi * 0.166667f));
            else ustep = tustep * (1.0f - gexp(-xi)) / xi;
            callhowfar = !(ustep < tperp);
        }
    }
    uint32_t irnew = p.region;
    int idisc = 0;
    const float ustep0 = ustep;
    if (callhowfar || p.wt <= 0.0f) irnew = howfar_module(env, p, ustep);
    if (irnew == 0) idisc = (rd.med == VACUUM || p.leaf == PASSMLC) ? 1 : -1;
    if (idisc > 0) {
        p.region = 0;
        return false;
    }
    if (ustep <= 0.0f) ustep = 0.0f;
    if (ustep == 0.0f || medium == (int)VACUUM) {
        if (ustep != 0.0f) {
            vstep = ustep;
            p.x = p.x + p.u * vstep;
            p.y = p.y + p.v * vstep;
            p.z = p.z + p.w * vstep;
            p.dnear = p.dnear - vstep;
        }
        if (p.region != irnew) t.gathers++;
        p.region = irnew;
        rd = region_rec(env, irnew);
        if (eie <= rd.ecut) {
            t.score(p.region, to_kinetic(p.e) * p.wt);
            if (p.charge > 0) annih_at_rest(env, s, p, rd, sink, t);
            return false;
        }
        if (ustep != 0.0f && idisc < 0) {
            t.score(p.region, (p.e + (float)p.charge * RM) * p.wt);
            return false;
        }
        is_tstep = true;
        return true;
    }
    vstep = ustep;
    if (callhowfar && exact_bca) {
        // single-scattering mode: a crossed boundary ends the step there and
        // shuts off the single scattering
        tvstep = vstep;
        if (tvstep != tustep) dosingle = false;
        de = eloss_over(mt, dedx, qel, medium, range, rhof, tvstep, eke, elke, lelke);
    } else if (callhowfar) {
        if (vstep < ustep0) {
            const float ekems = eke - 0.5f * tustep * vstep / ustep0 * dedx_val;
            const float p2m = ekems * (ekems + RMT2);
            beta2 = p2m / (p2m + RMSQ);
            chia2 = xccl / (4.0f * blccl * p2m);
            xi = 0.5f * xccl / p2m / beta2 * vstep;
            xi = xi * (glog(1.0f + 1.0f / chia2) - 1.0f / (1.0f + chia2));
            if (xi < 0.1f) {
                tvstep = vstep * (1.0f + xi * (0.5f + xi * 0.333333f));
            } else if (xi < 0.999999f) {
                tvstep = -vstep * glog(1.0f - xi) / xi;
            } else {
                return false;
            }
        } else {
            tvstep = tustep;
        }
        de = eloss_over(mt, dedx, qel, medium, range, rhof, tvstep, eke, elke, lelke);
    } else {
        tvstep = tustep;
        if (!callmsdist) de = eloss_over(mt, dedx, qel, medium, range, rhof, tvstep, eke, elke, lelke);
    }
    const float save_de = de;
    if (!callmsdist) {
        float sinthe = 0.0f, costhe = 1.0f;
        if (domultiple) {
            const float lambda = blccl * tvstep / beta2 / (1.0f + chia2);
            mscat(mt, s, lambda, chia2, xi, costhe, sinthe);
        } else if (dosingle) {
            // single scattering at the end of the step (electron.cu:2247-2276)
            const float ekems = fmaxf(eke - de, rd.ecut - RM);
            const float p2s = ekems * (ekems + RMT2);
            chia2 = mt->xcc[medium] / (4.0f * mt->blcc[medium] * p2s);
            sscat(s, chia2, costhe, sinthe);
        }
        p.x = p.x + p.u * vstep;
        p.y = p.y + p.v * vstep;
        p.z = p.z + p.w * vstep;
        if (domultiple || dosingle) {
            float sinphi, cosphi;
            uphi21(s, costhe, sinthe, cosphi, sinphi, p);
        }
    }
    p.dnear = p.dnear - vstep;
    const uint32_t irold = p.region;
    eie = eie - de;
    p.e = eie;
    t.score(irold, de * p.wt);
    if (irnew == irl && eie <= rd.ecut) {
        t.score(p.region, to_kinetic(p.e) * p.wt);
        if (p.charge > 0) annih_at_rest(env, s, p, rd, sink, t);
        return false;
    }
    const int medold = medium;
    if (medium != (int)VACUUM) {
        eke = eie - RM;
        elke = glog(eke);
        const gegs_float2 ek = mt->eke01[medium];
        lelke = (int)(ek.x + ek.y * elke);
    }
    if (irnew != irold) {
        p.region = irnew;
        rd = region_rec(env, irnew);
        medium = rd.med;
        t.gathers++;
    }
    if (eie <= rd.ecut) {
        t.score(p.region, to_kinetic(p.e) * p.wt);
        if (p.charge > 0) annih_at_rest(env, s, p, rd, sink, t);
        return false;
    }
    if (idisc < 0) {
        t.score(p.region, (p.e + (float)p.charge * RM) * p.wt);
        return false;
    }
    if (medium != medold) {
        is_tstep = true;
        return true;
    }
    demfp = demfp - save_de * sig_val;
    total_tstep = total_tstep - tvstep * rhof;
    if (total_tstep < 1e-9f) demfp = 0.0f;
    if ((double)demfp >= EPSEMFP) {
        is_tstep = false;
        return true;
    }
    const gegs_float2 sd = mt->sig[qel * MXEKE * MXMED + medium * MXEKE + lelke - 1];
    const gegs_float2 dd = mt->dedx[qel * MXEKE * MXMED + medium * MXEKE + lelke - 1];
    float sigf = sd.x + sd.y * elke;
    sigf /= dd.x + dd.y * elke;
    const float sigratio = sigf / sig0;
    if (s.next() > sigratio) {
        is_tstep = true;
        return true;
    }
    if (lelec < 0) {
        const gegs_float2 b1 = mt->ebr1[medium * MXEKE + lelke - 1];
        const float ebr1 = b1.x + b1.y * elke;
        if (s.next() <= ebr1) { forward(sink, s, p, E_BREMS); return false; }
        if (p.e <= mt->thmoll[medium] && mt->eii_flag == 0) {
            if (ebr1 <= 0.0f) { is_tstep = true; return true; }
            forward(sink, s, p, E_BREMS);
            return false;
        }
        forward(sink, s, p, E_MOLLER);
        return false;
    }
    const gegs_float2 pb1 = mt->pbr1[medium * MXEKE + lelke - 1];
    const float pbr1 = pb1.x + pb1.y * elke;
    const float r25 = s.next();
    if (r25 < pbr1) { forward(sink, s, p, E_BREMS); return false; }
    const gegs_float2 pb2 = mt->pbr2[medium * MXEKE + lelke - 1];
    const float pbr2 = pb2.x + pb2.y * elke;
    forward(sink, s, p, r25 < pbr2 ? E_BHABHA : E_ANNIH);
    return false;
}

template <bool VOTE = true, class Sink, class T>
__device__ __forceinline__ void electron_step(const Env &env, Stream &s, Part &p, Sink &sink, T &t) {
    EState st;
    st.init();
    for (;;) {
        if (VOTE) wave_lookahead(s, true);
        if (!electron_iter(env, s, p, st, sink, t)) break;
    }
}

// ------------------------------------------------------------------ electron-impact ionisation
// eii_sample (eii_sample.cu:4-96): the energy transfer to an electron of
// shell ish (binding uj) of element iz from the mixture of distant and close
// collisions, the secondary created above AE, then the vacancy relaxed;
// edep is set, not accumulated, as the reference does
template <class Sink>
__device__ __forceinline__ void eii_sample(const Env &env, Stream &s, int ish, int iz, float uj, Part &p,
                                           const gegs_region &rd, float &edep, Sink &sink) {
    const gegs_media_tables *mt = env.mt;
    Part p1 = p;
    const float eie = p.e;
    const float t = eie - RM;
    const float tau = t / RM;
    const float tau12 = (tau + 1) * (tau + 1);
    const float p2 = tau * tau + 2 * tau;
    const float beta2 = p2 / tau12;
    const float wmax = (float)(0.5 * (double)(t + uj));
    const float xmax = uj / wmax;
    const float c1 = (wmax / eie) * (wmax / eie);
    const float c2 = (2 * tau + 1) / tau12;
    const float fm_s = glog(RMT2 * p2 / uj) - beta2 - 0.5f;
    const float prob_s = (float)(0.66666667 * (double)fm_s * (double)(1 + xmax + xmax * xmax));
    float fm_h = 2 + c1 - c2;
    if (fm_h < 1) fm_h = 1;
    const float prob = fm_h + prob_s;
    float r1, frej, wx;
    do {
        r1 = s.next();
        if ((r1 * prob) < fm_h) {
            r1 = s.next();
            wx = 1 / (r1 * xmax + 1 - r1);
            const float wxx = wx * xmax;
            const float aux = wxx / (2 - wxx);
            frej = (1 + aux * (aux - c2) + c1 * wxx * wxx) / fm_h;
        } else {
            r1 = s.next();
            // powf(x, 0.333333333f) through the library's own log and exp
            wx = 1.0f / gexp(glog(r1 * xmax * xmax * xmax + 1 - r1) * 0.333333333f);
            frej = 1 - glog(wx) / fm_s;
        }
        r1 = s.next();
    } while (r1 >= frej);
    wx = wx * uj;
    const float h1 = (eie + RM) / t;
    const float ese1 = eie - wx;
    p.e = ese1;
    float dcosth = h1 * (ese1 - RM) / (ese1 + RM);
    float sinthe = sqrtf(1 - dcosth);
    float costhe = sqrtf(dcosth);
    float sinphi, cosphi;
    uphi21(s, costhe, sinthe, cosphi, sinphi, p);
    const float ese2 = wx - uj + RM;
    if (ese2 > mt->ae[rd.med]) {
        p1.e = ese2;
        dcosth = h1 * (ese2 - RM) / (ese2 + RM);
        sinthe = -sqrtf(1 - dcosth);
        costhe = sqrtf(dcosth);
        turn(costhe, sinthe, cosphi, sinphi, p1);
        create(sink, s, p1, E_ELECTRON_STEP);
        edep = 0;
    } else {
        edep = wx - uj;
    }
    relax(env, s, uj, ish, iz, edep, p, rd, sink);
}

// the EII branch of moller (moller.cu:78-121): true when the interaction was
// an ionisation of an EII shell (then handled completely)
template <class Sink, class T>
__device__ __forceinline__ bool eii_select(const Env &env, Stream &s, Part &p, const gegs_region &rd, Sink &sink, T &t) {
    const gegs_media_tables *mt = env.mt;
    const int m = rd.med;
    const float ekin = p.e - RM;
    const float elke = glog(ekin);
    const uint32_t lelke = (uint32_t)(mt->eke01[m].y * elke + mt->eke01[m].x);
    const gegs_float2 sg = mt->sig[m * MXEKE + lelke - 1];
    const gegs_float2 eb = mt->ebr1[m * MXEKE + lelke - 1];
    float sigm = sg.y * elke + sg.x;
    const float pbrem = eb.y * elke + eb.x;
    sigm = sigm * (1 - pbrem);
    float rsh = s.next();
    rsh = sigm * rsh;
    const int nne = mt->nne[m];
    for (int iele = 0; iele < nne; iele++) {
        const int iz = (int)((double)mt->zelem[iele * MXMED + m] + 0.5);
        const int nsh = mt->eii_no[iele * MXMED + m];
        if (nsh <= 0) continue;
        const int ifirst = mt->eii_first[iele * MXMED + m];
        for (int ish = 1; ish <= nsh; ish++) {
            const float uj = mt->binding_energies[(iz - 1) * MXSHELL + ish - 1];
            if ((ekin > uj) && (uj > mt->te[m] || uj > mt->ap[m])) {
                const int jj = ifirst + ish - 1;
                const int i = (int)(mt->eii[jj - 1].x * elke + mt->eii[jj - 1].y + (float)((jj - 1) * N_EII_BINS));
                const gegs_float2 xs = mt->eii_xsection[i - 1];
                float sig_j = xs.x * elke + xs.y;
                sig_j = sig_j * mt->pz[MXMED * iele + m] * mt->eii_cons[m];
                rsh -= sig_j;
                if (rsh < 0.0f) {
                    float edep = 0.0f;
                    eii_sample(env, s, ish, iz, uj, p, rd, edep, sink);
                    if (p.e < rd.ecut) edep += p.e - RM; else create(sink, s, p, E_ELECTRON_STEP);
                    t.score(p.region, edep * p.wt);
                    return true;
                }
            }
        }
    }
    return false;
}

// ------------------------------------------------------------------ Moller
template <class Sink, class T>
__device__ __forceinline__ void moller(const Env &env, Stream &s, Part &p, Sink &sink, T &t) {
    const gegs_media_tables *mt = env.mt;
    const gegs_region rd = region_rec(env, p.region);
    const int m = rd.med;
    float edep = 0.0f;
    Part p1 = p;
    const float eie = p.e;
    const float ekin = to_kinetic(eie);
    const float te = mt->te[m];
    if (mt->eii_flag > 0 && mt->eii_nsh[m] > 0 && eii_select(env, s, p, rd, sink, t)) return;
    if (ekin <= 2.0f * te) {
        t.score(p.region, to_kinetic(p.e) * p.wt);
        return;
    }
    const float t0 = ekin / RM;
    const float e0 = t0 + 1.0f;
    const float extrae = eie - mt->thmoll[m];
    const float e02 = e0 * e0;
    const float g2 = t0 * t0 / e02;
    const float g3 = (2.0f * t0 + 1.0f) / e02;
    const float gmax = 1.0f + 1.25f * g2;
    float r, rejf4, br, rr;
    do {
        r = s.next();
        br = te / (ekin - extrae * r);
        rr = br / (1.0f - br);
        r = s.next();
        rejf4 = 1.0f + g2 * br * br + rr * (rr - g3);
        r = gmax * r;
    } while (r > rejf4);
    const float ekse2 = br * ekin;
    const float ese1 = eie - ekse2;
    const float ese2 = ekse2 + RM;
    p.e = ese1;
    p1.e = ese2;
    const float h1 = (eie + RM) / ekin;
    float dcosth = h1 * (ese1 - RM) / (ese1 + RM);
    float sinthe = sqrtf(1.0f - dcosth);
    float costhe = sqrtf(dcosth);
    float sinphi, cosphi;
    uphi21(s, costhe, sinthe, cosphi, sinphi, p);
    if (p.e < rd.ecut) edep += to_kinetic(p.e); else create(sink, s, p, E_ELECTRON_STEP);
    dcosth = h1 * (ese2 - RM) / (ese2 + RM);
    sinthe = -sqrtf(1.0f - dcosth);
    costhe = sqrtf(dcosth);
    turn(costhe, sinthe, cosphi, sinphi, p1);
    if (p1.e < rd.ecut) edep += to_kinetic(p1.e); else create(sink, s, p1, E_ELECTRON_STEP);
    t.score(p.region, edep * p.wt);
}

// ------------------------------------------------------------------ Bhabha
template <class Sink, class T>
__device__ __forceinline__ void bhabha(const Env &env, Stream &s, Part &p, Sink &sink, T &t) {
    const gegs_media_tables *mt = env.mt;
    const gegs_region rd = region_rec(env, p.region);
    float edep = 0.0f;
    Part p1 = p;
    const float eip = p.e;
    const float ekin = to_kinetic(eip);
    const float t0 = ekin / RM;
    const float e0 = t0 + 1.0f;
    const float yy = 1.0f / (t0 + 2.0f);
    const float e02 = e0 * e0;
    const float beta2 = (e02 - 1.0f) / e02;
    const float ep0 = mt->te[rd.med] / ekin;
    const float ep0c = 1.0f - ep0;
    const float y2 = yy * yy;
    const float yp = 1.0f - 2.0f * yy;
    const float yp2 = yp * yp;
    const float b4 = yp2 * yp;
    const float b3 = b4 + yp2;
    const float b2 = yp * (3.0f + y2);
    const float b1 = 2.0f - y2;
    float r, rejf2, br;
    do {
        r = s.next();
        br = ep0 / (1.0f - ep0c * r);
        r = s.next();
        rejf2 = 1.0f - beta2 * br * (b1 - br * (b2 - br * (b3 - br * b4)));
    } while (r > rejf2);
    if (br < 0.5f) {
        p1.charge = -1;
    } else {
        p.charge = -1;
        p1.charge = 1;
        br = 1.0f - br;
    }
    br = fmaxf(br, 0.0f);
    const float ekse2 = br * ekin;
    const float ese1 = eip - ekse2;
    const float ese2 = ekse2 + RM;
    p.e = ese1;
    p1.e = ese2;
    const float h1 = (eip + RM) / ekin;
    float dcosth = fminf(1.0f, h1 * (ese1 - RM) / (ese1 + RM));
    float sinthe = sqrtf(1.0f - dcosth);
    float costhe = sqrtf(dcosth);
    float sinphi, cosphi;
    uphi21(s, costhe, sinthe, cosphi, sinphi, p);
    if (p.e < rd.ecut) {
        edep += to_kinetic(p.e);
        if (p.charge > 0) annih_at_rest(env, s, p, rd, sink, t);
    } else {
        create(sink, s, p, E_ELECTRON_STEP);
    }
    dcosth = h1 * (ese2 - RM) / (ese2 + RM);
    sinthe = -sqrtf(1.0f - dcosth);
    costhe = sqrtf(dcosth);
    turn(costhe, sinthe, cosphi, sinphi, p1);
    if (p1.e < rd.ecut) {
        edep += to_kinetic(p1.e);
        if (p1.charge > 0) annih_at_rest(env, s, p1, rd, sink, t);
    } else {
        create(sink, s, p1, E_ELECTRON_STEP);
    }
    t.score(p.region, edep * p.wt);
}

// ------------------------------------------------------------------ bremsstrahlung (Bethe-Heitler + Koch-Motz 2BS)
template <class Sink, class T>
__device__ __forceinline__ void brems(const Env &env, Stream &s, Part &p, Sink &sink, T &t) {
    const gegs_media_tables *mt = env.mt;
    const gegs_region rd = region_rec(env, p.region);
    const int m = rd.med;
    const int m8 = 8 * m;
    const int nsplit = mt->nbr_split;
    float edep = 0.0f;
    Part g = p;
    if (nsplit < 1) return;
    const float eie = p.e;
    const int l = eie < 50.0f ? 1 : 3;
    const float ekin = to_kinetic(eie);
    const float elke = glog(ekin);
    const float brmin = mt->ap[m] / ekin;
    const float waux = elke - mt->log_ap[m];
    const float a = p.u, b = p.v, cc = p.w;
    float sinpsi = a * a + b * b, sindel = 0.0f, cosdel = 0.0f;
    if (sinpsi > 1e-20f) {
        sinpsi = sqrtf(sinpsi);
        sindel = b / sinpsi;
        cosdel = a / sinpsi;
    }
    const float ztarg = mt->zbrang[m];
    const float tteie = eie / RM;
    const float beta = sqrtf((tteie - 1.0f) * (tteie + 1.0f)) / tteie;
    const float y2max = 2.0f * beta * (1.0f + beta) * tteie * tteie;
    const float y2maxi = 1.0f / y2max;
    const float z2maxi = sqrtf(y2max + 1.0f);
    float ese = eie;
    for (int ib = 1; ib <= nsplit; ++ib) {
        float r, rejf, esg;
        do {
            r = s.next();
            const float br = brmin * gexp(r * waux);
            esg = ekin * br;
            ese = eie - esg;
            const float delta = esg / eie / ese * mt->delcm[m];
            const float aux = ese / eie;
            float phi1, phi2;
            if (delta < 1.0f) {
                phi1 = mt->dl1[m8 + l - 1] + delta * (mt->dl2[m8 + l - 1] + delta * mt->dl3[m8 + l - 1]);
                phi2 = mt->dl1[m8 + l] + delta * (mt->dl2[m8 + l] + delta * mt->dl3[m8 + l]);
            } else {
                phi1 = mt->dl4[m8 + l - 1] + mt->dl5[m8 + l - 1] * glog(delta + mt->dl6[m8 + l - 1]);
                phi2 = phi1;
            }
            rejf = (1.0f + aux * aux) * phi1 - 2.0f * aux * phi2 / 3.0f;
            r = s.next();
        } while (r >= rejf);
        g.e = esg;
        g.charge = 0;
        g.wt = p.wt / (float)nsplit;
        float y2tst;
        {
            const float esedei = ese / RM / tteie;
            const float rj1 = 1.0f + esedei * esedei;
            const float rj2 = rj1 + 2.0f * esedei;
            float aux = 2.0f * ese * tteie / esg;
            aux = aux * aux;
            const float aux1 = aux * ztarg;
            const float rj3 = aux1 > 10.0f ? mt->lzbrang[m] + (1.0f - aux1) / (aux1 * aux1) : glog(aux / (1.0f + aux1));
            const float rejmax = rj1 * rj3 - rj2;
            float rtest, rejtst;
            do {
                y2tst = s.next();
                rtest = s.next();
                const float aux3 = z2maxi / (y2tst + (1.0f - y2tst) * z2maxi);
                rtest = rtest * aux3 * rejmax;
                y2tst = aux3 * aux3 - 1.0f;
                const float a4 = aux3 * aux3 * aux3 * aux3;
                const float y2tst1 = esedei * y2tst / a4;
                const float aux4 = 16.0f * y2tst1 - rj2;
                const float aux5 = rj1 - 4.0f * y2tst1;
                if (rtest < aux4 + aux5 * rj3) break;
                const float aux2 = glog(aux / (1.0f + aux1 / a4));
                rejtst = aux4 + aux5 * aux2;
            } while (rtest >= rejtst);
        }
        const float costhe = 1.0f - 2.0f * y2tst * y2maxi;
        const float sinthe = sqrtf(fmaxf((1.0f - costhe) * (1.0f + costhe), 0.0f));
        float cphi, sphi;
        azimuth_box(s, cphi, sphi);
        if (sinpsi >= 1e-10f) {
            const float us = sinthe * cphi, vs = sinthe * sphi;
            g.u = cc * cosdel * us - sindel * vs + a * costhe;
            g.v = cc * sindel * us + cosdel * vs + b * costhe;
            g.w = cc * costhe - sinpsi * us;
        } else {
            g.u = sinthe * cphi;
            g.v = sinthe * sphi;
            g.w = cc * costhe;
        }
        if (g.e < rd.pcut) edep += g.e; else create(sink, s, g, P_PHOTON_STEP);
    }
    p.e = ese;
    if (p.e < rd.ecut) edep += to_kinetic(p.e); else create(sink, s, p, E_ELECTRON_STEP);
    t.score(p.region, edep * p.wt);
}

// ------------------------------------------------------------------ in-flight annihilation (Heitler)
template <class Sink, class T>
__device__ __forceinline__ void annih(const Env &env, Stream &s, Part &p, Sink &sink, T &t) {
    const gegs_media_tables *mt = env.mt;
    const gegs_region rd = region_rec(env, p.region);
    const int nsplit = mt->nbr_split;
    float edep = 0.0f;
    if (nsplit <= 0) return;
    const float avip = p.e + RM;
    const float a = avip / RM;
    const float g = a - 1.0f;
    const float tt = g - 1.0f;
    const float pp = sqrtf(a * tt);
    const float pot = pp / tt;
    const float ep0 = 1.0f / (a + pp);
    const float wsamp = glog((1.0f - ep0) / ep0);
    const float aa = p.u, bb = p.v, cc = p.w;
    float sinpsi = aa * aa + bb * bb, sindel = 0.0f, cosdel = 0.0f;
    if (sinpsi > 1e-20f) {
        sinpsi = sqrtf(sinpsi);
        sindel = bb / sinpsi;
        cosdel = aa / sinpsi;
    }
    if (nsplit > 1) p.wt = p.wt / (float)nsplit;
    for (int ib = 1; ib <= nsplit; ++ib) {
        float r, rejf, ep;
        do {
            r = s.next();
            ep = ep0 * gexp(r * wsamp);
            r = s.next();
            rejf = 1.0f - (ep * a - 1.0f) * (ep * a - 1.0f) / (ep * (a * a - 2.0f));
        } while (r > rejf);
        const float esg1 = avip * ep;
        p.e = esg1;
        p.charge = 0;
        float costhe = fminf(1.0f, (esg1 - RM) * pot / esg1);
        float sinthe = sqrtf(1.0f - costhe * costhe);
        float cphi, sphi;
        azimuth_box(s, cphi, sphi);
        for (int k = 0; k < 2; ++k) {
            if (k == 1) {
                const float esg2 = avip - esg1;
                p.e = esg2;
                costhe = fminf(1.0f, (esg2 - RM) * pot / esg2);
                sinthe = -sqrtf(1.0f - costhe * costhe);
            }
            if (sinpsi >= 1e-10f) {
                const float us = sinthe * cphi, vs = sinthe * sphi;
                p.u = cc * cosdel * us - sindel * vs + aa * costhe;
                p.v = cc * sindel * us + cosdel * vs + bb * costhe;
                p.w = cc * costhe - sinpsi * us;
            } else {
                p.u = sinthe * cphi;
                p.v = sinthe * sphi;
                p.w = cc * costhe;
            }
            if (p.e < rd.pcut) edep += p.e; else create(sink, s, p, P_PHOTON_STEP);
        }
    }
    t.score(p.region, edep * p.wt);
}

// dispatch one particle to its event handler
template <bool VOTE = true, class Sink, class T>
__device__ __forceinline__ void handle(const Env &env, Stream &s, Part &p, Sink &sink, T &t) {
    switch (p.status) {
    case P_PHOTON_STEP: photon_step<VOTE>(env, s, p, sink, t); break;
    case P_COMPTON: compton(env, s, p, sink, t); break;
    case P_PHOTO: photo(env, s, p, sink, t); break;
    case P_RAYLEIGH: rayleigh(env, s, p, sink, t); break;
    case P_PAIR: pair(env, s, p, sink, t); break;
    case E_ELECTRON_STEP: electron_step<VOTE>(env, s, p, sink, t); break;
    case E_MOLLER: moller(env, s, p, sink, t); break;
    case E_BREMS: brems(env, s, p, sink, t); break;
    case E_BHABHA: bhabha(env, s, p, sink, t); break;
    case E_ANNIH: annih(env, s, p, sink, t); break;
    default: break;
    }
}

}  // namespace gegs
