/*
 * egs_oracle.c -- CPU restatement of the reference's coupled e-/photon
 * transport, used as the correctness oracle and as the CPU baseline.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this code, and only as the checker; the product
 * path (gegs_amd/) never calls it.
 *
 * One history at a time with its own particle stack (the "plain" method of
 * the small water case), LIFO, processed particle by particle.  Each function
 * follows the reference file cited in its comment; the reference's ten global
 * stacks (CUDA_EGS.h:362-367) become one per-history stack whose records keep
 * the event type in `status`, so a push "to the compton stack" is a push with
 * status = compton.
 *
 * Random numbers: Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) instead
 * of cuRAND XORWOW (random.cu).  Every particle owns a counter-based stream
 * keyed by (seed, lineage); a stream is buffered four uniforms at a time only
 * while its particle is being processed, and a particle stored on the stack
 * keeps just its next block index.  Particles created by an interaction get a
 * lineage derived from the parent's lineage and stream position, so the
 * random numbers every particle sees do not depend on the order in which
 * particles are processed.  Uniforms are in (0, 1] like curand_uniform.
 *
 * Arithmetic: binary32 with explicit float literals; quantities the
 * reference keeps in double (prm, the FP64 tallies) stay double.
 * Transcendentals come from oracle_math.h.  Compile with -ffp-contract=off
 * and without -ffast-math.
 *
 * Physics switches follow the reference defaults (CUDA_EGS.h:34-51): PRESTA-I,
 * inexact BCA, no spin, KM bremsstrahlung angles, IPRDST=1, EII on,
 * Rayleigh on, photo-electron angles on.  Bound Compton and atomic
 * relaxation are controlled per region by the region flags.
 *
 * Reference quirks kept ("reference-compat"): Q1 hownear_phantom uses
 * region-5 (geom.cu:1487-1490); its read past the last bound is defined here
 * as one extrapolated bound (the bounds arrays carry N+2 entries).  Q3 the
 * Compton br loop cap never fires (compton.cu:134-141).  Q5 nbr_split <= 0
 * ends annih.  Q6 photo's low-energy test against edge_energies[1].  Q7 range
 * rejection compares total energy with e_max_rr (electron.cu:1523).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/gegs.h"
#include "oracle_math.h"

#define RM 0.5110034f
static const double PRM = 0.5110034;
#define RMT2 (2.0f * RM)
#define RMSQ (RM * RM)
#define VACUUM_STEP 1e8f
#define EPSGMFP 1e-5
#define EPSEMFP 1e-5
#define EPSTEP 0.0005f
#define INFI_DIST 1e5f
#define VAC_AND_OTH 6
#define MXMED GEGS_MXMED
#define MXGE GEGS_MXGE
#define MXEKE GEGS_MXEKE
#define MXRAYFF GEGS_MXRAYFF
#define RAYCDFSIZE GEGS_RAYCDFSIZE
#define MAXL_MS GEGS_MAXL_MS
#define MAXQ_MS GEGS_MAXQ_MS
#define MAXU_MS GEGS_MAXU_MS
#define LAMBMAX_MS 1e5f
#define QMIN_MS 1e-3f
#define QMAX_MS 0.5f
#define HC_INVERSE 80.65506856998f
#define TWICE_HC2 0.000307444456f
#define STACK_MAX 64

enum {
    P_PHOTON_STEP = 0, P_COMPTON = 1, P_PHOTO = 2, P_RAYLEIGH = 3, P_PAIR = 4,
    E_ELECTRON_STEP = 5, E_MOLLER = 6, E_BREMS = 7, E_BHABHA = 8, E_ANNIH = 9
};

typedef struct opart {
    float e, x, y, z, u, v, w, wt, dnear;
    uint32_t region;
    int32_t charge;
    int32_t status;
    int32_t module;
    uint64_t lin;
    uint32_t blk;
    int32_t leaf;      /* MLC leaf index, OUTMLC / PASSMLC outside the leaf rows */
    int32_t bank;      /* MLC bank: 0 opening, 1 A, 2 B, -1 not located */
} opart;

#define OUTMLC (-99)
#define PASSMLC (-98)
#define NLEAF GEGS_MLC_LEAVES

/* ------------------------------------------------------------ Philox */
static void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

static uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct ostream {
    uint64_t lin;
    uint32_t blk;
    int pos;
    int nchild;
    float buf[4];
    uint32_t k0, k1;
} ostream;

static void st_begin(ostream *s, const opart *p, uint64_t seed) {
    s->lin = p->lin;
    s->blk = p->blk;
    s->pos = 4;
    s->nchild = 0;
    s->k0 = (uint32_t)seed;
    s->k1 = (uint32_t)(seed >> 32);
}

static float rnd(ostream *s) {
    if (s->pos == 4) {
        uint32_t c[4] = {s->blk, (uint32_t)s->lin, (uint32_t)(s->lin >> 32), 0x47454753u};
        philox(c, s->k0, s->k1);
        for (int i = 0; i < 4; ++i) s->buf[i] = (float)((c[i] >> 8) + 1u) * 5.96046448e-08f;
        s->blk += 1;
        s->pos = 0;
    }
    return s->buf[s->pos++];
}

static uint64_t child_lineage(ostream *s) {
    uint64_t pos = (uint64_t)s->blk * 4u + (uint64_t)s->pos;
    uint64_t k = (uint64_t)(s->nchild++);
    return mix64(s->lin ^ mix64((pos << 8) | k));
}

/* ------------------------------------------------------------ context */
typedef struct octx {
    const gegs_media_tables *mt;
    const gegs_region *reg;
    const float *xb, *yb, *zb;      /* N+2 entries each */
    uint32_t nx, ny, nz;
    gegs_field f;
    const gegs_mlc *mlc;            /* NULL: no MLC */
    const gegs_float2 *endnodes;    /* [60][34] leaf ends of the field */
    const gegs_float2 *jaws;        /* [32] jaw polygons of the field (NULL: no jaws) */
    int jawx_next;                  /* module after the X jaws */
    const gegs_block_wedge *bw;     /* block and wedge (NULL: neither) */
    int mlc_next, block_next;       /* modules after the MLC and after the block */
    uint32_t fixes;                 /* GEGS_FIX_* corrections (oracle_set_fixes) */
    uint64_t seed;
    double *dose, *dose2, *hist_edep;
    uint32_t *hist_steps;
    double *tmp;                     /* per-voxel deposit of the current history */
    uint64_t *last;                  /* history tag of tmp */
    uint64_t hist;
    double cur_edep;
    uint32_t cur_steps;
    double escaped;                  /* energy leaving the phantom or scored outside it */
    double dropped;                  /* energy of particles the transport abandons */
    double path;                     /* summed electron path length (tvstep) of the handler call */
    opart stack[STACK_MAX];
    int sp;
    gegs_counters cnt;
} octx;

/* GEGS_FIX_* mask that every context made after oracle_set_fixes uses */
static uint32_t g_fixes = 0;
void oracle_set_fixes(uint32_t mask) { g_fixes = mask; }

static void push(octx *c, const opart *p) {
    if (c->sp >= STACK_MAX) {
        c->cnt.overflow++;
        c->dropped += ((double)p->e + (double)p->charge * PRM) * (double)p->wt;
        return;
    }
    c->stack[c->sp++] = *p;
}

/* push the particle under processing to the next event (same stream) */
static void push_cont(octx *c, ostream *s, opart *p, int status) {
    p->status = status;
    p->lin = s->lin;
    p->blk = s->blk;
    push(c, p);
}

/* push a particle created by the current interaction */
static void push_child(octx *c, ostream *s, opart *p, int status) {
    p->status = status;
    p->lin = child_lineage(s);
    p->blk = 0;
    push(c, p);
}

/* escore, scoring.cu:14-24: FP64 tally, regions below 6 are not scored.
 * History-by-history variance: per-voxel tag of the last history. */
static void escore(octx *c, uint32_t region, float energy) {
    if (region < VAC_AND_OTH) { c->escaped += (double)energy; return; }
    double tmp = energy;
    uint32_t v = region - VAC_AND_OTH;
    c->cnt.deposits++;
    c->dose[v] += tmp;
    if (c->last[v] != c->hist + 1) {
        if (c->last[v] != 0) c->dose2[v] += c->tmp[v] * c->tmp[v];
        c->tmp[v] = 0.0;
        c->last[v] = c->hist + 1;
    }
    c->tmp[v] += tmp;
    c->cur_edep += tmp;
}

/* ------------------------------------------------------------ geometry */
static void to_phantom(const octx *c, const opart *p0, float *q) {
    const gegs_field *f = &c->f;
    if (c->fixes & GEGS_FIX_FP64_FRAME) {
        /* binary64 transform: the cos/sin pair normalised to unit length, one
         * rounding per coordinate */
        double cs = (double)f->cos_angle, sn = (double)f->sin_angle;
        const double r = sqrt(cs * cs + sn * sn);
        cs = cs / r;
        sn = sn / r;
        double x, y, z, u, v, w;
        if (f->change_xyz) {
            x = p0->y; y = (double)p0->z - (double)f->source_to_isocenter; z = p0->x;
            u = p0->v; v = p0->w; w = p0->u;
        } else {
            x = p0->x; y = p0->y; z = (double)p0->z - (double)f->source_to_isocenter;
            u = p0->u; v = p0->v; w = p0->w;
        }
        q[0] = (float)(x * cs - y * sn + (double)f->iso_x);
        q[1] = (float)(x * sn + y * cs + (double)f->iso_y);
        q[2] = (float)(z + (double)f->iso_z);
        q[3] = (float)(u * cs - v * sn);
        q[4] = (float)(u * sn + v * cs);
        q[5] = (float)w;
        return;
    }
    float x, y, z, u, v, w;
    if (f->change_xyz) {
        x = p0->y; y = p0->z - f->source_to_isocenter; z = p0->x;
        u = p0->v; v = p0->w; w = p0->u;
    } else {
        x = p0->x; y = p0->y; z = p0->z - f->source_to_isocenter;
        u = p0->u; v = p0->v; w = p0->w;
    }
    float tx = x * f->cos_angle - y * f->sin_angle;
    float ty = x * f->sin_angle + y * f->cos_angle;
    float tu = u * f->cos_angle - v * f->sin_angle;
    float tv = u * f->sin_angle + v * f->cos_angle;
    q[0] = tx + f->iso_x; q[1] = ty + f->iso_y; q[2] = z + f->iso_z;
    q[3] = tu; q[4] = tv; q[5] = w;
}

/* isWhere, geom.cu:20-36 */
static int is_where(float p, uint32_t nreg, const float *b) {
    if (p < b[0] || p > b[nreg]) return -1;
    if (nreg == 1) return 0;
    int ml = 0, mu = (int)nreg;
    while (mu - ml > 1) {
        int mav = (ml + mu) / 2;
        if (p <= b[mav]) mu = mav; else ml = mav;
    }
    return mu - 1;
}

/* howfar_phantom, geom.cu:54-307 */
static uint32_t howfar_phantom(const octx *c, const opart *p0, float *t) {
    float q[6];
    to_phantom(c, p0, q);
    const float px = q[0], py = q[1], pz = q[2], pu = q[3], pv = q[4], pw = q[5];
    const uint32_t nx = c->nx, ny = c->ny, nz = c->nz;
    const float *xb = c->xb, *yb = c->yb, *zb = c->zb;
    if (p0->region >= VAC_AND_OTH) {
        uint32_t ir = p0->region - VAC_AND_OTH;
        int iz = (int)(ir / (nx * ny));
        ir -= (uint32_t)iz * nx * ny;
        int iy = (int)(ir / nx);
        int ix = (int)(ir - (uint32_t)iy * nx);
        uint32_t inew = p0->region;
        if (pu > 0.0f) {
            float d = (xb[ix + 1] - px) / pu;
            if (d <= *t) { *t = d; if (++ix < (int)nx) inew = p0->region + 1; else inew = 0; }
        } else if (pu < 0.0f) {
            float d = (xb[ix] - px) / pu;
            if (d <= *t) { *t = d; if (--ix >= 0) inew = p0->region - 1; else inew = 0; }
        }
        if (pv > 0.0f) {
            float d = (yb[iy + 1] - py) / pv;
            if (d <= *t) { *t = d; if (++iy < (int)ny) inew = p0->region + nx; else inew = 0; }
        } else if (pv < 0.0f) {
            float d = (yb[iy] - py) / pv;
            if (d <= *t) { *t = d; if (--iy >= 0) inew = p0->region - nx; else inew = 0; }
        }
        if (pw > 0.0f) {
            float d = (zb[iz + 1] - pz) / pw;
            if (d <= *t) { *t = d; if (++iz < (int)nz) inew = p0->region + nx * ny; else inew = 0; }
        } else if (pw < 0.0f) {
            float d = (zb[iz] - pz) / pw;
            if (d <= *t) { *t = d; if (--iz >= 0) inew = p0->region - nx * ny; else inew = 0; }
        }
        if (*t < 0.0f) *t = EPSTEP;
        return inew;
    }
    int ix, iy, iz;
    float t1 = 0.0f;
    ix = -1;
    if (px <= xb[0] && pu > 0.0f) { t1 = (xb[0] - px) / pu; ix = 0; }
    else if (px >= xb[nx] && pu < 0.0f) { t1 = (xb[nx] - px) / pu; ix = (int)nx - 1; }
    if (ix >= 0 && t1 <= *t) {
        float y1 = py + pv * t1;
        iy = is_where(y1, ny, yb);
        if (iy >= 0) {
            float z1 = pz + pw * t1;
            iz = is_where(z1, nz, zb);
            if (iz >= 0) { *t = t1; return (uint32_t)(ix + iy * (int)nx + iz * (int)(nx * ny)) + VAC_AND_OTH; }
        }
    }
    iy = -1;
    if (py <= yb[0] && pv > 0.0f) { t1 = (yb[0] - py) / pv; iy = 0; }
    else if (py >= yb[ny] && pv < 0.0f) { t1 = (yb[ny] - py) / pv; iy = (int)ny - 1; }
    if (iy >= 0 && t1 <= *t) {
        float x1 = px + pu * t1;
        ix = is_where(x1, nx, xb);
        if (ix >= 0) {
            float z1 = pz + pw * t1;
            iz = is_where(z1, nz, zb);
            if (iz >= 0) { *t = t1; return (uint32_t)(ix + iy * (int)nx + iz * (int)(nx * ny)) + VAC_AND_OTH; }
        }
    }
    iz = -1;
    if (pz <= zb[0] && pw > 0.0f) { t1 = (zb[0] - pz) / pw; iz = 0; }
    else if (pz >= zb[nz] && pw < 0.0f) { t1 = (zb[nz] - pz) / pw; iz = (int)nz - 1; }
    if (iz >= 0 && t1 <= *t) {
        float x1 = px + pu * t1;
        ix = is_where(x1, nx, xb);
        if (ix >= 0) {
            float y1 = py + pv * t1;
            iy = is_where(y1, ny, yb);
            if (iy >= 0) { *t = t1; return (uint32_t)(ix + iy * (int)nx + iz * (int)(nx * ny)) + VAC_AND_OTH; }
        }
    }
    return 0;
}

/* hownear_phantom, geom.cu:1422-1543, with quirk Q1 (region - 5); with
 * GEGS_FIX_Q1 set by oracle_set_fixes, the particle's own voxel (region - 6) */
static float hownear_phantom(const octx *c, const opart *p0) {
    float q[6];
    to_phantom(c, p0, q);
    const float px = q[0], py = q[1], pz = q[2];
    const uint32_t nx = c->nx, ny = c->ny, nz = c->nz;
    const float *xb = c->xb, *yb = c->yb, *zb = c->zb;
    float tperp = INFI_DIST;
    const uint32_t first = (c->fixes & GEGS_FIX_Q1) ? 6u : 5u;
    if (p0->region >= first) {
        uint32_t ir = p0->region - first;
        int iz = (int)(ir / (nx * ny));
        ir -= (uint32_t)iz * nx * ny;
        int iy = (int)(ir / nx);
        int ix = (int)(ir - (uint32_t)iy * nx);
        float a, tx, ty, tz;
        a = xb[ix + 1] - px; tx = px - xb[ix]; if (a < tx) tx = a;
        a = yb[iy + 1] - py; ty = py - yb[iy]; if (a < ty) ty = a;
        a = zb[iz + 1] - pz; tz = pz - zb[iz]; if (a < tz) tz = a;
        if (tx < ty) tperp = tx; else tperp = ty;
        if (tperp > tz) tperp = tz;
    } else {
        float p1x = xb[0], p1y = yb[0], p1z = zb[0];
        float p2x = xb[nx], p2y = yb[ny], p2z = zb[nz];
        if (px - p1x > 0 && px - p2x > 0) p2x = p2x;
        else if (px - p1x < 0 && px - p2x < 0) p2x = p1x;
        else p2x = px;
        if (py - p1y > 0 && py - p2y > 0) p2y = p2y;
        else if (py - p1y < 0 && py - p2y < 0) p2y = p1y;
        else p2y = py;
        if (pz - p1z > 0 && pz - p2z > 0) p2z = p2z;
        else if (pz - p1z < 0 && pz - p2z < 0) p2z = p1z;
        else p2z = pz;
        float dx = px - p2x, dy = py - p2y, dz = pz - p2z;
        tperp = sqrtf(dx * dx + dy * dy + dz * dz);
    }
    return tperp;
}


/* ------------------------------------------------------------ MLC (geom.cu:310-790, :1152-1286) */
typedef struct { float x, y; } pt2;

static pt2 mk2(float x, float y) { pt2 q; q.x = x; q.y = y; return q; }

/* crossing number of a ray towards -x (geom.cu:310-327) */
static int in_poly(pt2 p, const gegs_float2 *v, int n) {
    int odd = 0, j = n - 1;
    for (int i = 0; i < n; ++i) {
        if (((v[i].y < p.y && v[j].y >= p.y) || (v[j].y < p.y && v[i].y >= p.y)) && (v[i].x <= p.x || v[j].x <= p.x)) {
            if (v[i].x + (p.y - v[i].y) / (v[j].y - v[i].y) * (v[j].x - v[i].x) < p.x) odd = !odd;
        }
        j = i;
    }
    return odd;
}

static int leaf_bank(pt2 p, const gegs_float2 *en) {
    if (in_poly(p, en, 17)) return 1;
    if (in_poly(p, en + 17, 17)) return 2;
    return 0;
}

/* leaf index of a point (x across the leaves, y = z), geom.cu:346-410 */
static int leaf_index(const gegs_mlc *m, pt2 p) {
    float x = p.x / p.y * m->zmin;
    float z = p.y - m->zmin;
    if (z < 0.0f || z > m->zthick) return OUTMLC;
    if (x < m->x_start || x > m->x_end) return OUTMLC;
    int i = OUTMLC;
    if (x < m->x_start + m->wt) i = 0;
    else if (x < m->x_mid_1 + m->wt) i = (int)((x - m->x_start - m->wt) / (m->wl_full + m->air_gap));
    else if (x < m->x_mid_2 + m->wt) {
        float pr = m->wl_target + m->wl_isocenter + m->air_gap * 2.0f;
        i = (int)((x - m->x_mid_1 - m->wt) / pr);
        i = (x - m->x_mid_1 - m->wt - pr * (float)i < m->wl_target + m->air_gap) ? 10 + i * 2 : 11 + i * 2;
    } else if (x < m->x_end - m->wl_full) i = (int)((x - m->x_mid_2 - m->wt) / (m->wl_full + m->air_gap)) + 50;
    else if (x <= m->x_end) i = 58;
    if (i < 0) return OUTMLC;
    if (in_poly(p, m->cross_xz + i * 16, 14)) return i;
    if (i < NLEAF - 1) return in_poly(p, m->cross_xz + (i + 1) * 16, 14) ? i + 1 : -(i + 1);
    return OUTMLC;
}

/* nearest edge along v (in units of |v|) and inside flag (geom.cu:442-484) */
static float ray_poly(pt2 o, pt2 v, const gegs_float2 *q, int n, int *inside) {
    float best = INFI_DIST + 1.0f;
    int hits = 0;
    for (int i = 0; i < n; ++i) {
        int j = i + 1 < n ? i + 1 : 0;
        float den = v.x * (q[j].y - q[i].y) - v.y * (q[j].x - q[i].x);
        if (den != 0.0f) {
            float d = ((q[j].x - q[i].x) * (o.y - q[i].y) - (q[j].y - q[i].y) * (o.x - q[i].x)) / den;
            float u = ((q[i].x - o.x) * v.y - (q[i].y - o.y) * v.x) / den;
            if (d >= 0.0f && u >= 0.0f && u < 1.0f) {
                hits++;
                if (d < best) best = d;
            }
        }
    }
    *inside = hits & 1;
    return best;
}

static float edge_dist(pt2 p, gegs_float2 v, gegs_float2 w) {
    float l2 = (v.x - w.x) * (v.x - w.x) + (v.y - w.y) * (v.y - w.y);
    if (l2 == 0.0f) return sqrtf((p.x - v.x) * (p.x - v.x) + (p.y - v.y) * (p.y - v.y));
    float t = ((p.x - v.x) * (w.x - v.x) + (p.y - v.y) * (w.y - v.y)) / l2;
    if (t < 0.0f) return sqrtf((p.x - v.x) * (p.x - v.x) + (p.y - v.y) * (p.y - v.y));
    if (t > 1.0f) return sqrtf((p.x - w.x) * (p.x - w.x) + (p.y - w.y) * (p.y - w.y));
    float ax = v.x + t * (w.x - v.x), ay = v.y + t * (w.y - v.y);
    return sqrtf((p.x - ax) * (p.x - ax) + (p.y - ay) * (p.y - ay));
}

static float poly_near(pt2 p, const gegs_float2 *q, int n) {
    float best = 1e9f;
    for (int i = 0; i < n; ++i) {
        float d = edge_dist(p, q[i], q[i + 1 < n ? i + 1 : 0]);
        if (d < best) best = d;
    }
    return best;
}

static const gegs_float2 *bank_polygon(const gegs_float2 *en, int bank, int *n) {
    if (bank == 0) { *n = 30; return en + 2; }
    *n = 17;
    return bank == 2 ? en + 17 : en;
}

/* region of a point: 3 (MLC W) inside a leaf, 1 (air) otherwise (geom.cu:486-516) */
static uint32_t mlc_where(const gegs_mlc *m, const gegs_float2 *ends, float qx, float qy, float qz, int *index, int *bank) {
    uint32_t r = 1u;
    if (*index == OUTMLC) *index = leaf_index(m, mk2(qx, qz));
    if (*index >= 0) {
        *bank = leaf_bank(mk2(qy, qz), ends + *index * 34);
        r = *bank > 0 ? 3u : 1u;
    }
    return r;
}

/* howfar_MLC, geom.cu:518-790 */
static uint32_t howfar_mlc(const octx *c, opart *p, float *t) {
    const gegs_mlc *m = c->mlc;
    const gegs_float2 *ends = c->endnodes;
    uint32_t inew = p->region;
    int bank = p->bank, index = p->leaf;
    int lo = 0, hi = NLEAF, in1, in2;
    if (p->y < -m->rmax || p->y > m->rmax) return 0u;
    pt2 a = mk2(p->x, p->z), b = mk2(p->y, p->z), va = mk2(p->u, p->w), vb = mk2(p->v, p->w);
    float dxz = ray_poly(a, va, m->boundary_xz, 4, &in1);
    float dyz = ray_poly(b, vb, m->boundary_yz, 4, &in2);
    if (dxz >= INFI_DIST || dyz >= INFI_DIST) return 0u;
    float bound = dxz < dyz ? dxz : dyz;
    int jump = 0;
    if (in1 && in2) {
        pt2 e = mk2(a.x + dxz * va.x, a.y + (dxz - 0.001f) * va.y);
        lo = leaf_index(m, e);
        if (lo < 0) lo = -lo;
    } else {
        inew = 0u;
        if (bound >= INFI_DIST) return inew;
        if (*t > bound) {
            *t = bound + EPSTEP;
            jump = 1;
        } else {
            return inew;
        }
    }
    if (!jump) {
        index = p->leaf > -NLEAF ? p->leaf : leaf_index(m, a);
        if (index >= 0) {
            int n;
            dxz = ray_poly(a, va, m->cross_xz + index * 16, 14, &in1);
            const gegs_float2 *yz = bank_polygon(ends + index * 34, p->bank, &n);
            dyz = ray_poly(b, vb, yz, n, &in2);
            if (dyz > dxz) {
                if (va.x < 0.0f && index > 0) index = -index;
                else if (index < NLEAF - 1) index = -(index + 1);
            }
        } else if (index > -NLEAF) {
            int sx = 1, i;
            if (va.x < 0.0f) {
                sx = -1;
                if (lo > 0) lo -= 1;
                hi = -index + 1;
            } else {
                hi = lo;
                lo = -index - 1;
            }
            i = -index - sx;
            if (lo < 0) lo = 0;
            if (hi > NLEAF - 1) hi = NLEAF - 1;
            if (i < 0) i = 0;
            if (i > NLEAF - 1) i = NLEAF - 1;
            float hit;
            do {
                hit = ray_poly(a, va, m->cross_xz + i * 16, 14, &in1);
                i += sx;
            } while (i >= lo && i <= hi && hit >= INFI_DIST);
            if (hit < dxz) {
                dxz = hit;
                index = i - sx;
            } else {
                index = PASSMLC;
                p->module = c->mlc_next;
            }
        }
        float dm = dxz < dyz ? dxz : dyz;
        int out = 0;
        if (dm + EPSTEP >= bound) {
            dm = bound;
            out = 1;
        }
        if (*t > dm) {
            *t = dm + EPSTEP;
            if (out) {
                p->leaf = PASSMLC;
                p->module = c->mlc_next;
                return 1u;
            }
        } else {
            return inew;
        }
    }
    inew = mlc_where(m, ends, p->x + *t * p->u, p->y + *t * p->v, p->z + *t * p->w, &index, &bank);
    p->leaf = index;
    p->bank = bank;
    return inew;
}

/* hownear_MLC, geom.cu:1193-1286 */
static float hownear_mlc(const octx *c, opart *p) {
    const gegs_mlc *m = c->mlc;
    const gegs_float2 *ends = c->endnodes;
    pt2 a = mk2(p->x, p->z), b = mk2(p->y, p->z);
    float txz = INFI_DIST + 1.0f, tyz = INFI_DIST + 1.0f;
    int in1 = 1, in2 = 1;
    if (p->leaf < -NLEAF) {
        in1 = in_poly(a, m->boundary_xz, 4);
        in2 = in_poly(b, m->boundary_yz, 4);
    }
    float bxz = poly_near(a, m->boundary_xz, 4), byz = poly_near(b, m->boundary_yz, 4);
    float bound = bxz < byz ? bxz : byz;
    if (!(in1 && in2)) return bound;
    int index;
    if (p->leaf > PASSMLC) index = p->leaf;
    else {
        index = leaf_index(m, a);
        p->leaf = index;
    }
    if (index >= 0) {
        int n;
        txz = poly_near(a, m->cross_xz + index * 16, 14);
        if (p->bank < 0) p->bank = leaf_bank(b, ends + index * 34);
        const gegs_float2 *yz = bank_polygon(ends + index * 34, p->bank, &n);
        tyz = poly_near(b, yz, n);
    } else if (index > -NLEAF) {
        int i = -index;
        float u = poly_near(a, m->cross_xz + i * 16, 14);
        txz = poly_near(a, m->cross_xz + (i - 1) * 16, 14);
        if (u < txz) txz = u;
    }
    return txz < tyz ? txz : tyz;
}

/* ------------------------------------------------------------ secondary jaws
 * howfar_SecJawsY / _SecJawsX (geom.cu:793-1035) and hownear (geom.cu:1288-1374)
 * as built with ROTATE_SecJaws_XY (CUDA_EGS.h:25): across-jaw points are
 * (y, z) and (x, z); the Y pair's polygons are tested in (x, z), the X
 * pair's in (y, z); the entry test against the field rectangle uses (y, x).
 * Polygons of the field: [0,4) minY, [4,8) minX, [8,12) maxY, [12,16) maxX,
 * [16,24) geometry Y, [24,32) geometry X. */
static uint32_t howfar_jaws(const octx *c, opart *p, float *t, int is_y) {
    const gegs_float2 *bmin = c->jaws + (is_y ? 0 : 4);
    const gegs_float2 *bmax = c->jaws + (is_y ? 8 : 12);
    const gegs_float2 *g = c->jaws + (is_y ? 16 : 24);
    pt2 a = mk2(p->y, p->z), b = mk2(p->x, p->z), va = mk2(p->v, p->w), vb = mk2(p->u, p->w);
    int ina, inb, dummy;
    float dxz = ray_poly(a, va, bmax, 4, &ina);
    float dyz = ray_poly(b, vb, bmax, 4, &inb);
    if (dxz >= INFI_DIST || dyz >= INFI_DIST) return 0u;
    float bound = dxz < dyz ? dxz : dyz;
    pt2 pp = is_y ? b : a, vv = is_y ? vb : va;
    uint32_t r;
    if (ina && inb) {
        if (in_poly(pp, g + 2, 4)) {
            r = 1u;
            float d1 = ray_poly(pp, vv, g, 4, &dummy);
            float d2 = ray_poly(pp, vv, g + 4, 4, &dummy);
            if (d1 < INFI_DIST) {
                if (*t > d1) { r = 2u; *t = d1 + EPSTEP; }
            } else if (d2 < INFI_DIST) {
                if (*t > d2) { r = 2u; *t = d2 + EPSTEP; }
            } else if (*t > bound) {
                *t = bound + EPSTEP;
                p->module = is_y ? GEGS_M_SECJAWX : c->jawx_next;
            }
        } else {
            r = 2u;
            float d0 = ray_poly(pp, vv, g + 2, 4, &dummy);
            if (d0 < INFI_DIST) {
                /* geom.cu:992: the X pair steps by dis_yz here */
                /* the X pair steps by the (x, z) boundary distance (quirk Q8, geom.cu:992);
                 * GEGS_FIX_Q8 steps by the distance to the opening */
                if (*t > d0) { r = 1u; *t = ((is_y || (c->fixes & GEGS_FIX_Q8)) ? d0 : dyz) + EPSTEP; }
            } else if (*t > bound) {
                r = 0u;
                *t = bound + EPSTEP;
            }
        }
    } else {
        r = 1u;
        if (*t > bound) {
            *t = bound + EPSTEP;
            pt2 q = mk2(p->y + *t * p->v, p->x + *t * p->u);
            if (!in_poly(q, bmin, 4)) {
                r = 0u;
            } else {
                if (is_y) q.x = q.y;
                q.y = p->z + *t * p->w;
                r = in_poly(q, g + 2, 4) ? 1u : 2u;
            }
        }
    }
    return r;
}

static float hownear_jaws(const octx *c, const opart *p, int is_y) {
    const gegs_float2 *bmax = c->jaws + (is_y ? 8 : 12);
    const gegs_float2 *g = c->jaws + (is_y ? 16 : 24);
    pt2 a = mk2(p->y, p->z), b = mk2(p->x, p->z);
    int ina = in_poly(a, bmax, 4), inb = in_poly(b, bmax, 4);
    float ba = poly_near(a, bmax, 4), bb = poly_near(b, bmax, 4);
    if (!(ina && inb)) return ba < bb ? ba : bb;
    /* Y: polygons in (x, z), the other bound is that of (y, z); X: the reverse */
    pt2 pp = is_y ? b : a;
    float own = is_y ? bb : ba, other = is_y ? ba : bb;
    float d = poly_near(pp, g + 2, 4);
    if (!in_poly(pp, g + 2, 4) && d > own) d = own;
    return d < other ? d : other;
}

/* ------------------------------------------------------------ block and wedge
 * howfar_block / howfar_wedge (geom.cu:1037-1149) and their hownear
 * (geom.cu:1376-1412) as built with ROTATE_Block_XY and ROTATE_Wedge_XY
 * (CUDA_EGS.h:28-29): the first cross-section (blockXZ, goem_Wed) is tested
 * with (y, z) and (v, w), the second (blockYZ, boundWed) with (x, z), (u, w).
 * Outside the two prisms a particle steps onto the nearer one and counts as
 * inside; inside both it leaves at the nearer exit for the next module; a
 * ray that misses either prism passes to the next module at once. */
static uint32_t howfar_prisms(opart *p, float *t, const gegs_float2 *qa, int na, const gegs_float2 *qb,
                              uint32_t reg, int next) {
    int ina, inb;
    float da = ray_poly(mk2(p->y, p->z), mk2(p->v, p->w), qa, na, &ina);
    float db = ray_poly(mk2(p->x, p->z), mk2(p->u, p->w), qb, 4, &inb);
    if (da >= INFI_DIST || db >= INFI_DIST) {
        p->module = next;
        return 1u;
    }
    float bound = da < db ? da : db;
    if (ina && inb) {
        if (*t > bound) {
            *t = bound + EPSTEP;
            p->module = next;
            return 1u;
        }
        return reg;
    }
    if (*t > bound) {
        *t = bound + EPSTEP;
        return reg;
    }
    return 1u;
}

static float hownear_prisms(const opart *p, const gegs_float2 *qa, int na, const gegs_float2 *qb) {
    float a = poly_near(mk2(p->y, p->z), qa, na);
    float b = poly_near(mk2(p->x, p->z), qb, 4);
    return a < b ? a : b;
}

static uint32_t howfar(octx *c, opart *p, float *t) {
    if (p->module == GEGS_M_PHANTOM) return howfar_phantom(c, p, t);
    if (p->module == GEGS_M_MLC) return howfar_mlc(c, p, t);
    if (p->module == GEGS_M_BLOCK) return howfar_prisms(p, t, c->bw->block_xz, 4, c->bw->block_yz, 4u, c->block_next);
    if (p->module == GEGS_M_WEDGE)
        return howfar_prisms(p, t, c->bw->wedge, c->bw->wedge_nver, c->bw->wedge_bound, 5u, GEGS_M_PHANTOM);
    return howfar_jaws(c, p, t, p->module == GEGS_M_SECJAWY);
}
static float hownear(octx *c, opart *p) {
    if (p->module == GEGS_M_PHANTOM) return hownear_phantom(c, p);
    if (p->module == GEGS_M_MLC) return hownear_mlc(c, p);
    if (p->module == GEGS_M_BLOCK) return hownear_prisms(p, c->bw->block_xz, 4, c->bw->block_yz);
    if (p->module == GEGS_M_WEDGE) return hownear_prisms(p, c->bw->wedge, c->bw->wedge_nver, c->bw->wedge_bound);
    return hownear_jaws(c, p, p->module == GEGS_M_SECJAWY);
}

/* module order jaws -> MLC -> block -> wedge -> phantom over the loaded
 * modules (CUDA_EGS.cu:1279-1298) */
static void set_module_order(octx *c, const gegs_mlc *mlc, const gegs_block_wedge *bw) {
    c->fixes = g_fixes;
    c->bw = bw;
    c->block_next = bw && bw->has_wedge ? GEGS_M_WEDGE : GEGS_M_PHANTOM;
    c->mlc_next = bw && bw->has_block ? GEGS_M_BLOCK : c->block_next;
    c->jawx_next = mlc ? GEGS_M_MLC : c->mlc_next;
}


/* ------------------------------------------------------------ uphi (utility.cu:94-218) */
static void rotate(float costhe, float sinthe, float cosphi, float sinphi, opart *p) {
    float A = p->u, B = p->v, C = p->w;
    float sinps2 = A * A + B * B;
    if (sinps2 < 1e-20f) {
        p->u = sinthe * cosphi;
        p->v = sinthe * sinphi;
        p->w = C * costhe;
    } else {
        float sinpsi = sqrtf(sinps2);
        float us = sinthe * cosphi;
        float vs = sinthe * sinphi;
        float sindel = B / sinpsi;
        float cosdel = A / sinpsi;
        p->u = C * cosdel * us - sindel * vs + A * costhe;
        p->v = C * sindel * us + cosdel * vs + B * costhe;
        p->w = -sinpsi * us + C * costhe;
    }
}

static void uphi21(ostream *s, float costhe, float sinthe, float *cosphi, float *sinphi, opart *p) {
    float r1 = rnd(s);
    om_sincos2pi(r1, sinphi, cosphi);
    rotate(costhe, sinthe, *cosphi, *sinphi, p);
}

static void uphi32(float costhe, float sinthe, float cosphi, float sinphi, opart *p) {
    rotate(costhe, sinthe, cosphi, sinphi, p);
}

/* box method for the azimuth, e.g. electron.cu:793-806 */
static void select_azimuth(ostream *s, float *cphi, float *sphi) {
    float xphi, yphi, rhophi2;
    do {
        xphi = rnd(s);
        xphi = 2.0f * xphi - 1.0f;
        yphi = rnd(s);
        rhophi2 = xphi * xphi + yphi * yphi;
    } while (rhophi2 > 1.0f);
    rhophi2 = 1.0f / rhophi2;
    *cphi = (xphi * xphi - yphi * yphi) * rhophi2;
    *sphi = 2.0f * xphi * yphi * rhophi2;
}

/* ------------------------------------------------------------ annih_at_rest (annihatrest.cu:10-68) */
static void annih_at_rest(octx *c, ostream *s, opart *p, const gegs_region *rd) {
    const gegs_media_tables *mt = c->mt;
    float edep = 0.0f;
    if (mt->nbr_split > 1) p->wt = p->wt / (float)mt->nbr_split;
    p->e = RM;
    p->charge = 0;
    for (int ibr = 1; ibr <= mt->nbr_split; ++ibr) {
        float costhe = rnd(s);
        costhe = 2.0f * costhe - 1.0f;
        float sinthe = sqrtf(fmaxf(0.0f, (1.0f - costhe) * (1.0f + costhe)));
        float cphi, sphi;
        select_azimuth(s, &cphi, &sphi);
        p->u = sinthe * cphi;
        p->v = sinthe * sphi;
        p->w = costhe;
        if (p->e < rd->pcut) edep += p->e;
        else push_child(c, s, p, P_PHOTON_STEP);
        p->u = -p->u; p->v = -p->v; p->w = -p->w;
        if (p->e < rd->pcut) edep += p->e;
        else push_child(c, s, p, P_PHOTON_STEP);
    }
    escore(c, p->region, edep * p->wt);
}

/* ------------------------------------------------------------ photon_step (photon.cu:6-292) */
static void photon_step(octx *c, ostream *s, opart *p) {
    const gegs_media_tables *mt = c->mt;
    gegs_region rd = c->reg[p->region];
    for (;;) {
        c->cnt.photon_steps++;
        c->cur_steps++;
        rd = c->reg[p->region];
        if (p->e <= rd.pcut) { escore(c, p->region, p->e * p->wt); return; }
        if (p->wt <= 0.0f) return;
        float r1 = rnd(s);
        float dpmfp = -om_logf(r1);
        float gle = om_logf(p->e);
        int lgle = 0;
        float tstep, gmfp_val = 0.0f, cohfac = 0.0f;
        uint32_t old_medium = rd.med;
        if (rd.med == GEGS_VACUUM) {
            tstep = VACUUM_STEP;
        } else {
            gegs_float2 ge = mt->ge[rd.med];
            lgle = (int)(ge.x + ge.y * gle);
            gegs_float2 gm = mt->gmfp[rd.med * MXGE + lgle];
            float gmfpr0 = gm.x + gm.y * gle;
            gmfp_val = gmfpr0 / rd.rhof;
            if (rd.flags & GEGS_F_RAYLEIGH) {
                gegs_float2 ch = mt->cohe[rd.med * MXGE + lgle];
                cohfac = ch.x + ch.y * gle;
                gmfp_val *= cohfac;
            }
            tstep = gmfp_val * dpmfp;
        }
        uint32_t new_region = howfar(c, p, &tstep);
        int idisc = 0;
        if (new_region == 0) idisc = (rd.med == GEGS_VACUUM || p->leaf == PASSMLC) ? 1 : -1;
        if (idisc > 0) { c->escaped += (double)p->e * (double)p->wt; p->region = 0; return; }
        p->x += p->u * tstep;
        p->y += p->v * tstep;
        p->z += p->w * tstep;
        if (rd.med != GEGS_VACUUM) dpmfp = fmaxf(0.0f, dpmfp - tstep / gmfp_val);
        old_medium = rd.med;
        if (new_region != p->region) {
            p->region = new_region;
            rd = c->reg[new_region];
            c->cnt.region_gathers++;
        }
        if (p->e <= rd.pcut || idisc < 0) { escore(c, p->region, p->e * p->wt); return; }
        if (rd.med != old_medium || rd.med == GEGS_VACUUM || (double)dpmfp >= EPSGMFP) continue;
        if (rd.flags & GEGS_F_RAYLEIGH) {
            r1 = rnd(s);
            if (r1 < 1.0f - cohfac) { push_cont(c, s, p, P_RAYLEIGH); return; }
        }
        gegs_float2 b1 = mt->gbr1[rd.med * MXGE + lgle];
        gegs_float2 b2 = mt->gbr2[rd.med * MXGE + lgle];
        float gbr1 = b1.x + b1.y * gle;
        float gbr2 = b2.x + b2.y * gle;
        r1 = rnd(s);
        if (r1 <= gbr1 && p->e > 2.0f * RM) push_cont(c, s, p, P_PAIR);
        else if (r1 < gbr2) push_cont(c, s, p, P_COMPTON);
        else push_cont(c, s, p, P_PHOTO);
        return;
    }
}

/* ------------------------------------------------------------ relax (relax.cu:14-238) */
#define MXSHELL GEGS_MXSHELL
#define MXTRANS GEGS_MXTRANS
#define MXEDGE GEGS_MXEDGE
#define MXMDSH GEGS_MXMDSH
#define MXVAC GEGS_MXVAC
#define N_EII_BINS GEGS_N_EII_BINS
#define RELAX_CUTOFF 0.001
static const int first_transition[5] = {1, 20, 27, 33, 38};
static const int last_transition[5] = {19, 26, 32, 37, 39};
static const int final_state[39] = {4, 3, 5, 6, 202, 302, 402, 404, 403, 303, 502, 503, 504, 602, 603, 604,
                                    505, 605, 606, 13, 14, 5, 6, 505, 605, 606, 14, 5, 6, 505, 605, 606,
                                    5, 6, 505, 605, 606, 6, 606};

/* isotropic direction of a relaxation particle (relax.cu:62-90): cos(theta)
 * uniform, azimuth by the box method on the half disk */
static void relax_direction(ostream *s, opart *p) {
    float eta = rnd(s);
    eta = 2 * eta - 1;
    p->w = eta;
    eta = (1 - eta) * (1 + eta);
    if (eta > 1e-20f) {
        eta = sqrtf(eta);
        float xphi, yphi, rhophi2;
        do {
            xphi = rnd(s);
            xphi = 2 * xphi - 1;
            yphi = rnd(s);
            rhophi2 = xphi * xphi + yphi * yphi;
        } while (rhophi2 > 1);
        rhophi2 = 1 / rhophi2;
        p->u = eta * (xphi * xphi - yphi * yphi) * rhophi2;
        p->v = eta * 2 * xphi * yphi * rhophi2;
    } else {
        p->u = 0.0f;
        p->v = 0.0f;
        p->w = 1.0f;
    }
}

/* fill a vacancy in shell n (1 K .. 6 N) of element iz carrying `energy`:
 * fluorescence photons, Auger and Coster-Kronig electrons above the cuts are
 * created, the rest is added to *edep.  A cascade deeper than MXVAC
 * vacancies (an out-of-bounds write in the reference) deposits the vacancy
 * locally instead. */
static void relax(octx *c, ostream *s, float energy, int n, int iz, float *edep, opart p, const gegs_region *rd) {
    const gegs_media_tables *mt = c->mt;
    int vac[MXVAC];
    float ea[MXVAC];
    if (n < 1 || n > MXSHELL) return;
    float e_cut = fminf(rd->ecut - RM, rd->pcut);
    e_cut = ((double)e_cut > RELAX_CUTOFF) ? e_cut : (float)RELAX_CUTOFF;
    if (energy <= e_cut) { *edep += energy; return; }
    int nvac = 1;
    vac[0] = n;
    ea[0] = energy;
    const float *be = mt->binding_energies + MXSHELL * (iz - 1);
    for (;;) {
        int shell = vac[nvac - 1];
        float ei = ea[nvac - 1];
        nvac -= 1;
        if (ei <= e_cut) {
            *edep += ei;
            if (nvac > 0) continue;
            break;
        }
        if (shell == MXSHELL) {          /* N-shell vacancy: an Auger electron */
            p.e = ei + RM;
            p.charge = -1;
            relax_direction(s, &p);
            if (p.e < rd->ecut) *edep += p.e - RM;
            else push_child(c, s, &p, E_ELECTRON_STEP);
            if (nvac > 0) continue;
            break;
        }
        float eta = rnd(s);
        int k;
        for (k = first_transition[shell - 1]; k <= last_transition[shell - 1] - 1; k++) {
            eta = eta - mt->relaxation_prob[MXTRANS * (iz - 1) + k - 1];
            if (eta <= 0.0f) break;
        }
        int fin = final_state[k - 1];
        int iql;
        float elcut, ex;
        if (nvac + (fin < 100 ? 1 : 2) > MXVAC) {
            *edep += ei;
            if (nvac > 0) continue;
            break;
        }
        if (fin < 100) {
            if (fin < 10) {              /* fluorescence */
                iql = 0;
                elcut = rd->pcut;
            } else {                     /* Coster-Kronig */
                fin = fin - 10;
                iql = -1;
                elcut = rd->ecut - RM;
            }
            float ef = be[fin - 1];
            ex = ei - ef;
            nvac += 1;
            vac[nvac - 1] = fin;
            ea[nvac - 1] = ef;
        } else {                         /* Auger: two new vacancies */
            int f1 = fin / 100;
            int f2 = fin - f1 * 100;
            nvac += 1;
            vac[nvac - 1] = f1;
            ea[nvac - 1] = be[f1 - 1];
            nvac += 1;
            vac[nvac - 1] = f2;
            ea[nvac - 1] = be[f2 - 1];
            iql = -1;
            ex = ei - ea[nvac - 1] - ea[nvac - 2];
            elcut = rd->ecut - RM;
        }
        if (ex <= elcut) {
            *edep += ex;
        } else {
            p.charge = iql;
            p.e = iql == 0 ? ex : ex + RM;
            relax_direction(s, &p);
            if (iql == 0) {
                if (p.e < rd->pcut) *edep += p.e;
                else push_child(c, s, &p, P_PHOTON_STEP);
            } else {
                if (p.e < rd->ecut) *edep += p.e - RM;
                else push_child(c, s, &p, E_ELECTRON_STEP);
            }
        }
    }
}

/* ------------------------------------------------------------ compton (compton.cu:17-492) */
/* ibcmpl from the region flags (compton.cu:87-95): 0x2 bound Compton with
 * rejections returning the photon to the photon step, 0x6 without Doppler
 * broadening and binding, 0xA rejections resampled; anything else Klein-Nishina */
static int compton_mode(uint16_t flags) {
    switch (flags & 0x000E) {
    case 0x0002: return 1;
    case 0x0006: return 2;
    case 0x000A: return 3;
    default: return 0;
    }
}

static void compton(octx *c, ostream *s, opart *p) {
    const gegs_media_tables *mt = c->mt;
    gegs_region rd = c->reg[p->region];
    float edep = 0.0f;
    opart p1 = *p;
    int medium = rd.med;
    float peig = p->e;
    float ko = peig / RM;
    float broi = 1 + 2 * ko;
    int first_time = 1;
    int ibcmpl = compton_mode(rd.flags);
    /* declared inside the resample loop in the reference and read again on a
     * resample without being set: kept from the previous pass */
    float bro = 0.0f, br = 0.0f, temp = 0.0f, sinthe = 0.0f, costhe = 0.0f;
    float alph1 = 0.0f, alpha = 0.0f, rejmax = 0.0f;
    for (int compton_loops = 0;; compton_loops++) {
        if (compton_loops > 100) break;
        float uj = 0.0f, jo = 0.0f;
        int j = 0;
        if (ibcmpl > 0) {
            /* shell: Walker alias table over the medium's sub-shells */
            float rnno17 = rnd(s);
            rnno17 = 1 + rnno17 * mt->n_shell[medium];
            int i = (int)rnno17;
            if (rnno17 > mt->eno_array[medium * MXMDSH + i - 1]) i = mt->eno_atbin_array[medium * MXMDSH + i - 1];
            j = mt->shell_array[medium * MXMDSH + i - 1];
            uj = mt->be_array[j - 1];
            if (ko <= uj) {
                if (ibcmpl == 1) { push_cont(c, s, p, P_PHOTON_STEP); break; }
                continue;
            }
            jo = mt->jo_array[j - 1];
        }
        /* scattering angle from Klein-Nishina; br_loops is never incremented
         * in the reference (Q3), GEGS_FIX_Q3 applies its cap */
        int br_loops = 0;
        do {
            if ((c->fixes & GEGS_FIX_Q3) && br_loops++ > 100) break;
            if (ko > 2) {
                if (first_time) {
                    alph1 = om_logf(broi);
                    bro = 1 / broi;
                    float alph2 = ko * (broi + 1) * bro * bro;
                    alpha = alph1 + alph2;
                }
                float aux, rejf3, rnno19;
                int rnno19_loops = 0;
                do {
                    float rnno15 = rnd(s);
                    float rnno16 = rnd(s);
                    if ((rnno15 * alpha) < alph1) br = om_expf(alph1 * rnno16) * bro;
                    else br = sqrtf(rnno16 * broi * broi + (1 - rnno16)) * bro;
                    temp = (1 - br) / (ko * br);
                    sinthe = fmaxf(0.0f, temp * (2 - temp));
                    aux = 1 + br * br;
                    rejf3 = aux - br * sinthe;
                    rnno19 = rnd(s);
                    rnno19_loops++;
                    if (rnno19_loops > 100) break;
                } while (rnno19 * aux > rejf3);
            } else {
                if (first_time) {
                    bro = (float)(1.0 / (double)broi);
                    rejmax = broi + bro;
                }
                float rnno16, rejf3;
                int rnno16_loops = 0;
                do {
                    float rnno15 = rnd(s);
                    rnno16 = rnd(s);
                    br = bro + (1 - bro) * rnno15;
                    temp = (1 - br) / (ko * br);
                    sinthe = fmaxf(0.0f, temp * (2 - temp));
                    rejf3 = 1 + br * br - br * sinthe;
                    if (rnno16_loops++ > 100) break;
                } while (rnno16 * br * rejmax > rejf3);
            }
            first_time = 0;
        } while (br < bro || br > 1);
        costhe = 1 - temp;
        if (ibcmpl == 0) {
            uj = 0;
        } else {
            /* rejection by the limited range of p_z, then p_z itself (compton.cu:228-405) */
            float aux = ko * (ko - uj) * temp;
            float aux1 = 2 * aux + uj * uj;
            float pzmax = aux - uj;
            if ((pzmax < 0) && (pzmax * pzmax >= aux1)) {
                if (ibcmpl == 1) { push_cont(c, s, p, P_PHOTON_STEP); break; }
                continue;
            }
            pzmax = pzmax / sqrtf(aux1);
            float qc2 = 1 + br * br - 2 * br * costhe;
            float qc = sqrtf(qc2);
            float af, fmax1, fpz;
            int rejected = 0;
            if (pzmax > 1) {
                pzmax = 1;
                af = 0;
                fmax1 = 1;
                fpz = 1;
            } else {
                float aux3 = 1 + 2 * jo * fabsf(pzmax);
                float aux4 = (float)(0.5 * (double)(1 - aux3 * aux3));
                fpz = (float)(0.5 * (double)om_expf(aux4));
                af = qc * (1 + br * (br - costhe) / qc2);
                if (af < 0) {
                    if (pzmax > 0) fpz = 1.0f - fpz;
                    float eta_incoh = rnd(s);
                    if (eta_incoh > fpz) rejected = 1;
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
                    float eta_incoh = rnd(s);
                    if (eta_incoh * jo > fpz1) rejected = 1;
                }
            }
            if (rejected) {
                if (ibcmpl == 1) { push_cont(c, s, p, P_PHOTON_STEP); break; }
                continue;
            }
            float pz = 0.0f;
            for (int pz_loops = 0;; pz_loops++) {
                if (pz_loops > 100) break;
                if (ibcmpl != 2) {
                    float rnno18 = rnd(s);
                    rnno18 = rnno18 * fpz;
                    if (rnno18 < 0.5f) {
                        rnno18 = (double)(2 * rnno18) > 1e-30 ? 2 * rnno18 : (float)1e-30;
                        pz = (float)(0.5 * (double)(1 - sqrtf(1 - 2 * om_logf(rnno18))) / (double)jo);
                    } else {
                        rnno18 = 2 * (1 - rnno18);
                        pz = (float)(0.5 * (double)(sqrtf(1 - 2 * om_logf(rnno18)) - 1) / (double)jo);
                    }
                    if (fabsf(pz) > 1) continue;
                    if ((double)pz < 0.15) {
                        float frej;
                        if ((double)pz < -0.15) frej = (float)(1.0 - (double)af * 0.15);
                        else frej = 1 + af * pz;
                        float eta = rnd(s);
                        if (eta * fmax1 > frej) continue;
                    }
                } else {
                    pz = 0;
                    uj = 0;
                }
                break;
            }
            /* energy of the scattered photon */
            if ((double)fabsf(pz) < 0.01) {
                br = br * (1 + pz * (qc + (br * br - costhe) * pz));
            } else {
                aux = 1 - pz * pz * br * costhe;
                aux1 = 1 - pz * pz * br * br;
                float aux2 = qc2 - br * br * pz * pz * sinthe;
                if ((double)aux2 > 1e-10) br = br / aux1 * (aux + pz * sqrtf(aux2));
            }
            uj = uj * RM;
        }
        float pesg = br * peig;
        float pese = peig - pesg - uj + RM;
        sinthe = sqrtf(sinthe);
        p->e = pesg;
        float sinphi, cosphi;
        uphi21(s, costhe, sinthe, &cosphi, &sinphi, p);
        if (p->e < rd.pcut) edep += p->e;
        else push_child(c, s, p, P_PHOTON_STEP);
        float aux = 1 + br * br - 2 * br * costhe;
        if ((double)aux > 1e-8) {
            costhe = (1 - br * costhe) / sqrtf(aux);
            sinthe = (1 - costhe) * (1 + costhe);
            if (sinthe > 0) sinthe = -sqrtf(sinthe);
            else sinthe = 0;
        } else {
            costhe = 0.0f;
            sinthe = -1.0f;
        }
        p1.e = pese;
        p1.charge = -1;
        uphi32(costhe, sinthe, cosphi, sinphi, &p1);
        if (p1.e < rd.ecut) edep += p1.e - RM;
        else push_child(c, s, &p1, E_ELECTRON_STEP);
        if (ibcmpl == 1 || ibcmpl == 3) {
            if ((double)uj > RELAX_CUTOFF) relax(c, s, uj, mt->shn_array[j - 1], mt->iz_array[j - 1], &edep, *p, &rd);
            else edep += uj;
        }
        break;
    }
    escore(c, p->region, edep * p->wt);
}

/* ------------------------------------------------------------ photo (photo.cu:1-279) */
static void sauter_angle(ostream *s, opart *p) {
    /* Sauter photo-electron direction (photo.cu:197-244) */
    float beta = sqrtf((p->e - RM) * (p->e + RM)) / p->e;
    float gamma = p->e / RM;
    float alpha = 0.5f * gamma - 0.5f + 1.0f / gamma;
    float ratio = beta / alpha;
    float xi, rnpht2, sinth2, costhe;
    do {
        float rnpht = rnd(s);
        rnpht = 2.0f * rnpht - 1.0f;
        if (ratio <= 0.2f) {
            float fkappa = rnpht + 0.5f * ratio * (1.0f - rnpht) * (1.0f + rnpht);
            if (gamma < 100.0f) {
                costhe = (beta + fkappa) / (1.0f + beta * fkappa);
            } else {
                if (fkappa > 0.0f)
                    costhe = 1.0f - (1.0f - fkappa) * (gamma - 3.0f) /
                             (2.0f * (1.0f + fkappa) * (gamma - 1.0f) * (gamma - 1.0f) * (gamma - 1.0f));
                else
                    costhe = (beta + fkappa) / (1.0f + beta * fkappa);
            }
            xi = (1.0f + beta * fkappa) * gamma * gamma;
        } else {
            xi = gamma * gamma * (1.0f + alpha * (sqrtf(1.0f + ratio * (2.0f * rnpht + ratio)) - 1.0f));
            costhe = (1.0f - 1.0f / xi) / beta;
        }
        sinth2 = fmaxf(0.0f, (1.0f - costhe) * (1.0f + costhe));
        rnpht2 = rnd(s);
    } while (rnpht2 > 0.5f * (1.0f + gamma) * sinth2 * xi / gamma);
    float sinthe = sqrtf(sinth2);
    float sinphi, cosphi;
    uphi21(s, costhe, sinthe, &cosphi, &sinphi, p);
}

static void photo(octx *c, ostream *s, opart *p) {
    const gegs_media_tables *mt = c->mt;
    gegs_region rd = c->reg[p->region];
    float edep = 0.0f;
    int medium = rd.med;
    float peig = p->e;
    int iedgfl = rd.flags & GEGS_F_ATOMIC_RELAXATION;
    int iphter = rd.flags & GEGS_F_PHOTOELECTRON_ANGLE;
    if (peig < mt->edge_energies[1]) {   /* quirk Q6: the second edge of hydrogen */
        p->charge = -1;
        p->e = peig + RM;
        if (p->e < rd.ecut) {
            edep = p->e - RM;
            escore(c, p->region, edep * p->wt);
        } else {
            push_child(c, s, p, E_ELECTRON_STEP);
        }
        return;
    }
    float gle = om_logf(p->e);
    int do_relax = 0;
    float e_vac = 0.0f;
    int iz = 0, k = 0;
    if (iedgfl != 0) {
        /* the element (photo.cu:87-154), then the shell (photo.cu:160-183) */
        if (mt->nne[medium] == 1) {
            iz = (int)((double)mt->zelem[medium] + 0.5);
        } else {
            float probs[GEGS_MXEL];
            float aux = peig * peig;
            float aux1 = aux * peig;
            aux = aux * sqrtf(peig);
            float sigtot = 0;
            for (k = 0; k < mt->nne[medium]; k++) {
                iz = (int)((double)mt->zelem[MXMED * k + medium] + 0.5);
                const int e0 = MXEDGE * (iz - 1);
                float sigma;
                if (peig > mt->edge_energies[e0]) {
                    sigma = (mt->edge_a[e0] + mt->edge_b[e0] / peig + mt->edge_c[e0] / aux + mt->edge_d[e0] / aux1) / peig;
                } else {
                    int j;
                    for (j = 2; j <= mt->edge_number[iz - 1]; j++)
                        if (peig >= mt->edge_energies[e0 + j - 1]) break;
                    sigma = mt->edge_a[e0 + j - 1] +
                            gle * (mt->edge_b[e0 + j - 1] + gle * (mt->edge_c[e0 + j - 1] + gle * mt->edge_d[e0 + j - 1]));
                    sigma = om_expf(sigma);
                }
                sigma = sigma * mt->pz[MXMED * k + medium];
                sigtot = sigtot + sigma;
                probs[k] = sigma;
            }
            float br = rnd(s);
            br = br * sigtot;
            for (k = 0; k < mt->nne[medium]; k++) {
                br = br - probs[k];
                if (br <= 0) break;
            }
            /* past the last element only by rounding (an out-of-range read in the reference) */
            if (k == mt->nne[medium]) k = mt->nne[medium] - 1;
            iz = (int)((double)mt->zelem[MXMED * k + medium] + 0.5);
        }
        const float *be = mt->binding_energies + MXSHELL * (iz - 1);
        if (peig <= be[5]) {             /* below the N shell: absorbed locally */
            edep += peig;
        } else {
            float br = rnd(s);
            for (k = 1; k <= 5; k++) {
                if (peig > be[k - 1]) {
                    float tmp = mt->interaction_prob[MXSHELL * (iz - 1) + k - 1];
                    if (br < tmp) break;
                    br = (br - tmp) / (1 - tmp);
                }
            }
            e_vac = be[k - 1];
            p->e = peig - e_vac + RM;
            do_relax = 1;
            p->charge = -1;
        }
    } else {
        p->e = peig + RM;
        p->charge = -1;
    }
    if (p->charge == -1) {
        if (p->e > rd.ecut) {
            if (iphter) sauter_angle(s, p);
            push_child(c, s, p, E_ELECTRON_STEP);
        } else {
            edep += p->e - RM;
        }
    }
    if (do_relax) relax(c, s, e_vac, k, iz, &edep, *p, &rd);
    escore(c, p->region, edep * p->wt);
}

/* ------------------------------------------------------------ rayleigh (rayleigh.cu:4-157) */
static void rayleigh(octx *c, ostream *s, opart *p) {
    const gegs_media_tables *mt = c->mt;
    gegs_region rd = c->reg[p->region];
    float gle = om_logf(p->e);
    gegs_float2 ge = mt->ge[rd.med];
    int lgle = (int)(ge.x + ge.y * gle);
    gegs_float2 pm = mt->pmax[rd.med * MXGE + lgle];
    float pmax_val = pm.x + pm.y * gle;
    float xmax = HC_INVERSE * p->e;
    float costhe = 0.0f, costhe2 = 0.0f;
    for (;;) {
        float xv;
        for (;;) {
            float r1 = rnd(s);
            float temp = r1 * pmax_val;
            int dwi = RAYCDFSIZE - 1;
            int ibin = (int)(temp * (float)dwi);
            int ib = (int)mt->i_array[rd.med * RAYCDFSIZE + ibin] - 1;
            int next_ib = (int)mt->i_array[rd.med * RAYCDFSIZE + ibin + 1] - 1;
            if (next_ib > ib) {
                for (;;) {
                    gegs_rayleigh rdat = mt->rayleigh_data[rd.med * MXRAYFF + ib + 1];
                    if (temp < rdat.fcum || ib >= RAYCDFSIZE - 2) break;
                    ib++;
                }
            }
            gegs_rayleigh rdat = mt->rayleigh_data[rd.med * MXRAYFF + ib];
            temp = (temp - rdat.fcum) * rdat.c_array;
            xv = rdat.xgrid * om_expf(om_logf(1.0f + temp) * rdat.b_array);
            if (xv < xmax) break;
        }
        float r2 = rnd(s);
        xv = xv / p->e;
        costhe = 1.0f - TWICE_HC2 * xv * xv;
        costhe2 = costhe * costhe;
        if (2.0f * r2 < 1.0f + costhe2) break;
    }
    float sinthe = sqrtf(1.0f - costhe2);
    float cosphi, sinphi;
    uphi21(s, costhe, sinthe, &cosphi, &sinphi, p);
    if (p->e < rd.pcut) escore(c, p->region, p->e * p->wt);
    else push_child(c, s, p, P_PHOTON_STEP);
}

/* ------------------------------------------------------------ pair (pair.cu:14-431), iprdst = 1 */
static void pair(octx *c, ostream *s, opart *p) {
    const gegs_media_tables *mt = c->mt;
    gegs_region rd = c->reg[p->region];
    int medium = rd.med;
    float edep = 0.0f;
    opart p1 = *p;
    float eig = p->e;
    float ese1, ese2;
    int iq1, iq2;
    if (eig <= 2.1f) {
        float rnno30 = rnd(s);
        /* prm + 0.5*rnno30*(eig - 2*prm) with the double literal 0.5 (pair.cu:89) */
        ese2 = (float)((double)RM + 0.5 * (double)rnno30 * (double)(eig - 2.0f * RM));
        rnno30 = rnd(s);
        ese1 = eig - ese2;
        if (rnno30 < 0.5f) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
    } else {
        float amax, bmax, aux1;
        int l, l1;
        if (eig < 50.0f) {
            l = 5;
            l1 = l + 1;
            float delta = 4.0f * mt->delcm[medium] / eig;
            if (delta < 1.0f) {
                amax = mt->dl1[8 * medium + l - 1] + delta * (mt->dl2[8 * medium + l - 1] + delta * mt->dl3[8 * medium + l - 1]);
                bmax = mt->dl1[8 * medium + l1 - 1] + delta * (mt->dl2[8 * medium + l1 - 1] + delta * mt->dl3[8 * medium + l1 - 1]);
            } else {
                float aux2 = om_logf(delta + mt->dl6[8 * medium + l - 1]);
                amax = mt->dl4[8 * medium + l - 1] + mt->dl5[8 * medium + l - 1] * aux2;
                bmax = mt->dl4[8 * medium + l1 - 1] + mt->dl5[8 * medium + l1 - 1] * aux2;
            }
            aux1 = 1.0f - RMT2 / eig;
            aux1 = aux1 * aux1;
            aux1 = aux1 * amax / 3.0f;
            aux1 = aux1 / (bmax + aux1);
        } else {
            l = 7;
            amax = mt->dl1[8 * medium + l - 1];
            bmax = mt->dl1[8 * medium + l];
            aux1 = mt->bpar[2 * medium + 1] *
                   (1.0f - mt->bpar[((c->fixes & GEGS_FIX_Q2) ? 2 : 8) * medium] * RM / eig);   /* Q2 */
        }
        float del0 = eig * mt->delcm[medium];
        float eavail = eig - RMT2;
        float rnno30, rejmax, rejf, eminus;
        do {
            rnno30 = rnd(s);
            float rnno31 = rnd(s);
            float br;
            if (rnno30 > aux1) {
                br = 0.5f * rnno31;
                rejmax = bmax;
                l1 = l + 1;
            } else {
                rnno30 = rnd(s);
                rnno30 = fmaxf(rnno30, rnno31);
                rnno31 = rnd(s);
                float b1 = fmaxf(rnno30, rnno31);
                br = 0.5f * (1.0f - b1);
                rejmax = amax;
                l1 = l;
            }
            eminus = br * eavail + RM;
            float eplus = eig - eminus;
            float delta = del0 / (eminus * eplus);
            if (delta < 1.0f)
                rejf = mt->dl1[8 * medium + l1 - 1] + delta * (mt->dl2[8 * medium + l1 - 1] + delta * mt->dl3[8 * medium + l1 - 1]);
            else
                rejf = mt->dl4[8 * medium + l1 - 1] + mt->dl5[8 * medium + l1 - 1] * om_logf(delta + mt->dl6[8 * medium + l1 - 1]);
            rnno30 = rnd(s);
        } while (rnno30 * rejmax > rejf);
        ese2 = eminus;
        ese1 = eig - ese2;
        rnno30 = rnd(s);
        if (rnno30 < 0.5f) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
    }
    p->e = ese1;
    p1.e = ese2;
    p->charge = iq1;
    p1.charge = iq2;
    float cosphi = 0.0f, sinphi = 0.0f;
    for (int ichrg = 1; ichrg <= 2; ++ichrg) {
        float ese = ichrg == 1 ? ese1 : ese2;
        float pse = sqrtf(fmaxf(0.0f, (ese - RM) * (ese + RM)));
        float costhe = rnd(s);
        costhe = 1.0f - 2.0f * costhe;
        float sinthe = RM * sqrtf((1.0f - costhe) * (1.0f + costhe)) / (pse * costhe + ese);
        costhe = (ese * costhe + pse) / (pse * costhe + ese);
        if (ichrg == 1) {
            uphi21(s, costhe, sinthe, &cosphi, &sinphi, p);
            if (p->e < rd.ecut) {
                edep += (p->e - RM);
                if (p->charge > 0) annih_at_rest(c, s, p, &rd);
            } else {
                push_child(c, s, p, E_ELECTRON_STEP);
            }
        } else {
            sinthe = -sinthe;
            uphi32(costhe, sinthe, cosphi, sinphi, &p1);
            if (p1.e < rd.ecut) {
                edep += (p1.e - RM);
                if (p1.charge > 0) annih_at_rest(c, s, &p1, &rd);
            } else {
                push_child(c, s, &p1, E_ELECTRON_STEP);
            }
        }
    }
    escore(c, p->region, edep * p->wt);
}

/* ------------------------------------------------------------ mscat (electron.cu:77-242), spin off */
static void mscat(const gegs_media_tables *mt, ostream *s, float lambda, float chia2, float q1,
                  float *cost, float *sint) {
    if (lambda <= 13.8f) {
        float sprob = rnd(s);
        float explambda = om_expf(-lambda);
        if (sprob < explambda) { *cost = 1.0f; *sint = 0.0f; return; }
        float wsum = (1.0f + lambda) * explambda;
        if (sprob < wsum) {
            float xi = rnd(s);
            xi = 2.0f * chia2 * xi / (1.0f - xi + chia2);
            *cost = 1.0f - xi;
            *sint = sqrtf(xi * (2.0f - xi));
            return;
        }
        if (lambda <= 1.0f) {
            float wprob = explambda;
            wsum = explambda;
            *cost = 1.0f;
            *sint = 0.0f;
            int icount = 0;
            do {
                icount = icount + 1;
                if (icount > 20) break;
                wprob = wprob * lambda / (float)icount;
                wsum = wsum + wprob;
                float xi = rnd(s);
                xi = 2.0f * chia2 * xi / (1.0f - xi + chia2);
                float cosz = 1.0f - xi;
                float sinz = xi * (2.0f - xi);
                if (sinz > 1.0e-20f) {
                    sinz = sqrtf(sinz);
                    xi = rnd(s);
                    float sphi, cphi;
                    om_sincos2pi(xi, &sphi, &cphi);
                    *cost = *cost * cosz - *sint * sinz * cphi;
                    *sint = sqrtf(fmaxf(0.0f, (1.0f - *cost) * (1.0f + *cost)));
                }
            } while (wsum <= sprob);
            return;
        }
    }
    if (lambda <= LAMBMAX_MS) {
        float llmbda = om_logf(lambda);
        float ai = llmbda * mt->dllambi;
        int msi = (int)ai;
        ai = ai - (float)msi;
        float xi = rnd(s);
        if (xi < ai) msi = msi + 1;
        int msj;
        if (q1 < QMIN_MS) msj = 0;
        else if (q1 < QMAX_MS) {
            float aj = q1 * mt->dqmsi;
            msj = (int)aj;
            aj = aj - (float)msj;
            xi = rnd(s);
            if (xi < aj) msj = msj + 1;
        } else msj = MAXQ_MS;
        float omega2;
        if (llmbda < 2.2299f)
            omega2 = chia2 * (lambda + 4.0f) * (1.347006f + llmbda * (0.209364f - llmbda * (0.45525f - llmbda * (0.50142f - 0.081234f * llmbda))));
        else
            omega2 = chia2 * (lambda + 4.0f) * (-2.77164f + llmbda * (2.94874f - llmbda * (0.1535754f - llmbda * 0.00552888f)));
        xi = rnd(s);
        float ak = xi * (float)MAXU_MS;
        int msk = (int)ak;
        ak = ak - (float)msk;
        int base = (msk * (MAXQ_MS + 1) + msj) * (MAXL_MS + 1) + msi;
        if (ak > mt->wms[base]) msk = mt->ims[base];
        int itmp = (msk * (MAXQ_MS + 1) + msj) * (MAXL_MS + 1) + msi;
        float a = mt->fms[itmp];
        float u = mt->ums[itmp];
        float du = mt->ums[((msk + 1) * (MAXQ_MS + 1) + msj) * (MAXL_MS + 1) + msi] - u;
        xi = rnd(s);
        if (fabsf(a) < 0.2f) {
            float x1 = 0.5f * (1.0f - xi) * a;
            u = u + xi * du * (1.0f + x1 * (1.0f - xi * a));
        } else {
            u = u - du / a * (1.0f - sqrtf(1.0f + xi * a * (2.0f + a)));
        }
        xi = omega2 * u / (1.0f + 0.5f * omega2 - u);
        if (xi > 1.99999f) xi = 1.99999f;
        *cost = 1.0f - xi;
        *sint = sqrtf(xi * (2.0f - xi));
        return;
    }
    /* lambda above the table: the reference leaves the angle unset; take no
     * deflection (unreachable for the step sizes the skin depth allows) */
    *cost = 1.0f;
    *sint = 0.0f;
}

/* msdist_pI (electron.cu:618-882), spin off */
static float msdist_pI(const gegs_media_tables *mt, ostream *s, float eloss, float tustep, opart *p0,
                       const gegs_region *rd) {
    int medium = rd->med;
    float e0 = p0->e - RM;
    float blccc = mt->blcc[medium];
    float xcccc = mt->xcc[medium];
    float e = e0 - 0.5f * eloss;
    float p2 = e * (e + RMT2);
    float beta2 = p2 / (p2 + RMSQ);
    float lambda = tustep * rd->rhof * blccc / beta2;
    float factor = 1.0f / (1.0f + 0.9784671f * e);
    float epsilon = eloss / e0;
    epsilon = epsilon / (1.0f - 0.5f * epsilon);
    float temp = 0.25f * (1.0f - factor * (1.0f - 0.333333f * factor)) * epsilon * epsilon;
    lambda = lambda * (1.0f + temp);
    float chia2 = xcccc / p2 / (4.0f * blccc);
    lambda = lambda / (1.0f + chia2);
    float chilog = om_logf(1.0f + 1.0f / chia2);
    float q1 = 2.0f * chia2 * (chilog * (1.0f + chia2) - 1.0f);
    float xi = q1 * lambda;
    float ws, sint;
    mscat(mt, s, lambda, chia2, xi, &ws, &sint);
    float cphi, sphi;
    select_azimuth(s, &cphi, &sphi);
    float us = sint * cphi;
    float vs = sint * sphi;
    float z, r, r2, z2, r2max;
    if (xi < 0.1f) z = 1.0f - xi * (0.5f - xi * (0.166666667f - 0.041666667f * xi));
    else z = (1.0f - om_expf(-xi)) / xi;
    r = 0.5f * sint;
    r2 = r * r;
    z2 = z * z;
    r2max = 1.0f - z2;
    if (r2max < r2) { r2 = r2max; r = sqrtf(r2); }
    float ut = r * cphi, vt = r * sphi, wt = z;
    float ustep = sqrtf(z2 + r2) * tustep;
    float sint02 = p0->u * p0->u + p0->v * p0->v;
    if (sint02 > 1e-20f) {
        float sint0 = sqrtf(sint02);
        float cphi0 = p0->u / sint0;
        float sphi0 = p0->v / sint0;
        float u2p = p0->w * us + sint0 * ws;
        ws = p0->w * ws - sint0 * us;
        us = u2p * cphi0 - vs * sphi0;
        vs = u2p * sphi0 + vs * cphi0;
        u2p = p0->w * ut + sint0 * wt;
        wt = p0->w * wt - sint0 * ut;
        ut = u2p * cphi0 - vt * sphi0;
        vt = u2p * sphi0 + vt * cphi0;
    } else {
        wt = p0->w * wt;
        ws = p0->w * ws;
    }
    p0->x = p0->x + tustep * ut;
    p0->y = p0->y + tustep * vt;
    p0->z = p0->z + tustep * wt;
    p0->u = us;
    p0->v = vs;
    p0->w = ws;
    return ustep;
}

/* msdist_pII (electron.cu:282-614), spin off: the step is split in two
 * halves with one multiple-scattering angle each; the transport distance
 * combines the two directions with the random substep lengths b, c and the
 * energy-loss corrections of the PRESTA-II paper (Kawrakow & Bielajew 1998) */
static float msdist_pII(const gegs_media_tables *mt, ostream *s, float eloss, float tustep, opart *p0,
                        const gegs_region *rd) {
    int medium = rd->med;
    float e0 = p0->e - RM;
    float blccc = mt->blcc[medium];
    float xcccc = mt->xcc[medium];
    float e = e0 - 0.5f * eloss;
    float tau = e / RM;
    float tau2 = tau * tau;
    float epsilon = eloss / e0;
    float epsilonp = eloss / e;
    /* the reference's denominator (24 tau2 + 48 tau + 72), kept */
    e = e * (1.0f - epsilonp * epsilonp * (6.0f + 10.0f * tau + 5.0f * tau2) / (24.0f * tau2 + 48.0f * tau + 72.0f));
    float p2 = e * (e + RMT2);
    float beta2 = p2 / (p2 + RMSQ);
    float chia2 = xcccc / p2 / (4.0f * blccc);
    float lambda = 0.5f * tustep * rd->rhof * blccc / beta2;
    float temp = epsilonp / ((tau + 1.0f) * (tau + 2.0f));
    float temp1 = 0.166666f * (4.0f + tau * (6.0f + tau * (7.0f + tau * (4.0f + tau)))) * temp * temp;
    lambda = lambda * (1.0f - temp1);
    /* spin off: etap = xi_corr = gamma = ms_corr = 1 */
    float gamma = 1.0f;
    lambda = lambda / (1.0f + chia2);
    float chilog = om_logf(1.0f + 1.0f / chia2);
    float q1 = 2.0f * chia2 * (chilog * (1.0f + chia2) - 1.0f);
    gamma = 6.0f * chia2 * (1.0f + chia2) * (chilog * (1.0f + 2.0f * chia2) - 2.0f) / q1 * gamma;
    float xi = q1 * lambda;
    float w1, sint1, w2, sint2, cphi1, sphi1, cphi2, sphi2;
    mscat(mt, s, lambda, chia2, xi, &w1, &sint1);
    select_azimuth(s, &cphi1, &sphi1);
    mscat(mt, s, lambda, chia2, xi, &w2, &sint2);
    select_azimuth(s, &cphi2, &sphi2);
    /* final direction relative to motion along z */
    float u2 = sint2 * cphi2;
    float v2 = sint2 * sphi2;
    float u2p = w1 * u2 + sint1 * w2;
    float us = u2p * cphi1 - v2 * sphi1;
    float vs = u2p * sphi1 + v2 * cphi1;
    float ws = w1 * w2 - sint1 * u2;
    xi = 2.0f * xi;      /* xi was for half the step; xi_corr = 1 */
    float eta = rnd(s);
    eta = sqrtf(eta);    /* sampled from 2 eta d(eta) */
    float eta1 = 0.5f * (1.0f - eta);
    float delta = 0.9082483f - (0.1020621f - 0.0263747f * gamma) * xi;
    /* energy-loss corrections of the substep coefficients */
    temp1 = 2.0f + tau;
    temp = (2.0f + tau * temp1) / ((tau + 1.0f) * temp1);
    temp = temp - (tau + 1.0f) / ((tau + 2.0f) * (chilog * (1.0f + chia2) - 1.0f));
    temp = temp * epsilonp;
    temp1 = 1.0f - temp;
    delta = delta + 0.40824829f * (epsilon * (tau + 1.0f) /
                                   ((tau + 2.0f) * (chilog * (1.0f + chia2) - 1.0f) * (chilog * (1.0f + 2.0f * chia2) - 2.0f)) -
                                   0.25f * temp * temp);
    float b = eta * delta;
    float cc = eta * (1.0f - delta);
    float w1v2 = w1 * v2;
    float ut = b * sint1 * cphi1 + cc * (cphi1 * u2 - sphi1 * w1v2) + eta1 * us * temp1;
    float vt = b * sint1 * sphi1 + cc * (sphi1 * u2 + cphi1 * w1v2) + eta1 * vs * temp1;
    float wt = eta1 * (1.0f + temp) + b * w1 + cc * w2 + eta1 * ws * temp1;
    float ustep = tustep * sqrtf(ut * ut + vt * vt + wt * wt);
    float sint02 = p0->u * p0->u + p0->v * p0->v;
    if (sint02 > 1e-20f) {
        float sint0 = sqrtf(sint02);
        float cphi0 = p0->u / sint0;
        float sphi0 = p0->v / sint0;
        u2p = p0->w * us + sint0 * ws;
        ws = p0->w * ws - sint0 * us;
        us = u2p * cphi0 - vs * sphi0;
        vs = u2p * sphi0 + vs * cphi0;
        u2p = p0->w * ut + sint0 * wt;
        wt = p0->w * wt - sint0 * ut;
        ut = u2p * cphi0 - vt * sphi0;
        vt = u2p * sphi0 + vt * cphi0;
    } else {
        wt = p0->w * wt;
        ws = p0->w * ws;
    }
    p0->x = p0->x + tustep * ut;
    p0->y = p0->y + tustep * vt;
    p0->z = p0->z + tustep * wt;
    p0->u = us;
    p0->v = vs;
    p0->w = ws;
    return ustep;
}

/* sscat (electron.cu:245-279), spin off: one screened-Rutherford scattering */
static void sscat(ostream *s, float chia2, float *cost, float *sint) {
    float xi = rnd(s);
    xi = 2.0f * chia2 * xi / (1.0f - xi + chia2);
    *cost = 1.0f - xi;
    *sint = sqrtf(xi * (2.0f - xi));
}

/* $COMPUTE-ELOSS-G: energy loss over path `step` (electron.cu:1649-1701) */
static float compute_eloss(const gegs_media_tables *mt, int qel, int medium, float range, float rhof,
                           float step, float eke, float elke, int lelke) {
    float de;
    const gegs_float2 *dedx = mt->dedx + qel * MXEKE * MXMED + medium * MXEKE;
    const float *rep = mt->range_ep;
    float tuss = range - rep[(medium * MXEKE + lelke - 1) * 2 + qel] / rhof;
    if (tuss >= step) {
        gegs_float2 dd = dedx[lelke - 1];
        float dedxmid = dd.x + dd.y * elke;
        float aux = dd.y / dedxmid;
        de = dedxmid * step * rhof;
        float fedep = de / eke;
        de = de * (1.0f - 0.5f * fedep * aux * (1.0f - 0.333333f * fedep * (aux - 1.0f - 0.25f * fedep * (2.0f - aux * (4.0f - aux)))));
    } else {
        int lelktmp = lelke;
        tuss = (range - step) * rhof;
        if (tuss <= 0.0f) {
            de = eke - mt->te[medium] * 0.99f;
        } else {
            while (tuss < rep[(medium * MXEKE + lelktmp - 1) * 2 + qel]) lelktmp = lelktmp - 1;
            gegs_float2 ek = mt->eke01[medium];
            float elktmp = ((float)(lelktmp + 1) - ek.x) / ek.y;
            float eketmp = mt->e_array[medium * MXEKE + lelktmp];
            tuss = (rep[(medium * MXEKE + lelktmp) * 2 + qel] - tuss) / rhof;
            gegs_float2 dd = dedx[lelktmp - 1];
            float dedxmid = dd.x + dd.y * elktmp;
            float aux = dd.y / dedxmid;
            de = dedxmid * tuss * rhof;
            float fedep = de / eketmp;
            de = de * (1.0f - 0.5f * fedep * aux * (1.0f - 0.333333f * fedep * (aux - 1.0f - 0.25f * fedep * (2.0f - aux * (4.0f - aux)))));
            de = de + eke - eketmp;
        }
    }
    return de;
}

/* $COMPUTE-DRANGE: path from energy eke down to bin edge / ekef (electron.cu:1285-1360) */
static float drange(const gegs_float2 *dedx, float eke, float elke, float ekef_or_edge, float elkf, int lbin) {
    float fedep = 1.0f - ekef_or_edge / eke;
    float elktmp = 0.5f * (elke + elkf + 0.25f * fedep * fedep * (1.0f + fedep * (1.0f + 0.875f * fedep)));
    gegs_float2 dd = dedx[lbin - 1];
    float dedxmid = dd.x + dd.y * elktmp;
    dedxmid = 1.0f / dedxmid;
    float aux = dd.y * dedxmid;
    float temp = fedep / (2.0f - fedep);
    aux = aux * (1.0f + 2.0f * aux) * temp * temp / 6.0f;
    return fedep * eke * dedxmid * (1.0f + aux);
}

/* ------------------------------------------------------------ electr_step (electron.cu:888-2622) */
static void electron_step(octx *c, ostream *s, opart *p) {
    const gegs_media_tables *mt = c->mt;
    float elke = 0.0f, demfp = 0.0f, sig0 = 0.0f, total_tstep = 0.0f;
    int lelke = 0;
    int compute_tstep = 1;
    int is_tstep = 1;
    for (;;) {
        uint32_t irold = p->region;
        uint32_t irl = irold;
        gegs_region rd = c->reg[irl];
        int medium = rd.med;
        float rhof = rd.rhof;
        float edep1;
        int lelec = p->charge;
        int qel = (1 + lelec) / 2;
        float eie = p->e;
        c->cnt.electron_steps++;
        c->cur_steps++;
        if (eie <= rd.ecut) {
            edep1 = (p->e - RM);
            escore(c, p->region, edep1 * p->wt);
            if (p->charge > 0) annih_at_rest(c, s, p, &rd);
            return;
        }
        if (p->wt <= 0.0f) return;
        float eke = eie - RM;
        if (is_tstep) {
            compute_tstep = 1;
            if (medium != GEGS_VACUUM) {
                float rnne1 = rnd(s);
                if (rnne1 == 0.0f) rnne1 = 1.0e-30f;
                demfp = fmaxf(-om_logf(rnne1), (float)EPSEMFP);
                elke = om_logf(eke);
                gegs_float2 ek = mt->eke01[medium];
                lelke = (int)(ek.x + ek.y * elke);
                if (mt->sig_ismonotone[medium * 2 + qel]) {
                    gegs_float2 sd = mt->sig[qel * MXEKE * MXMED + medium * MXEKE + lelke - 1];
                    gegs_float2 dd = mt->dedx[qel * MXEKE * MXMED + medium * MXEKE + lelke - 1];
                    float sigf = sd.x + sd.y * elke;
                    float dedx0 = dd.x + dd.y * elke;
                    sigf /= dedx0;
                    sig0 = sigf;
                } else {
                    sig0 = mt->sig_e[qel * MXMED + medium];
                }
            }
        }
        float sig_val = 0.0f, dedx_val = 0.0f, tstep = 0.0f, ustep = 0.0f, tustep = 0.0f;
        float tvstep = 0.0f, vstep = 0.0f, range = 0.0f, de = 0.0f;
        float blccl = 0.0f, xccl = 0.0f, beta2 = 0.0f, chia2 = 0.0f, xi = 0.0f;
        float ekems, tperp = 0.0f;
        int callhowfar = 0, domultiple = 0, dosingle = 0, callmsdist = 0;
        const gegs_float2 *dedx = mt->dedx + qel * MXEKE * MXMED + (medium == GEGS_VACUUM ? 0 : medium) * MXEKE;
        if (medium == GEGS_VACUUM) {
            tstep = VACUUM_STEP;
            ustep = tstep;
            tustep = ustep;
            callhowfar = 1;
        } else {
            gegs_float2 ek = mt->eke01[medium];
            sig_val = sig0;
            if (sig_val <= 0.0f) {
                tstep = VACUUM_STEP;
                sig0 = 1.0e-15f;
            } else {
                if (compute_tstep) {
                    float total_de = demfp / sig_val;
                    float ekef = eke - total_de;
                    if (ekef <= mt->e_array[medium * MXEKE + 0]) {
                        tstep = VACUUM_STEP;
                    } else {
                        float elkef = om_logf(ekef);
                        int lelkef = (int)(ek.x + ek.y * elkef);
                        if (lelkef == lelke) {
                            tstep = drange(dedx, eke, elke, ekef, elkef, lelke);
                        } else {
                            float ekei = mt->e_array[medium * MXEKE + lelke - 1];
                            float elkei = ((float)lelke - ek.x) / ek.y;
                            float tuss = drange(dedx, eke, elke, ekei, elkei, lelke);
                            ekei = mt->e_array[medium * MXEKE + lelkef];
                            elkei = ((float)(lelkef + 1) - ek.x) / ek.y;
                            tstep = drange(dedx, ekei, elkei, ekef, elkef, lelkef);
                            tstep = tstep + tuss + mt->range_ep[(medium * MXEKE + lelke - 1) * 2 + qel]
                                    - mt->range_ep[(medium * MXEKE + lelkef) * 2 + qel];
                        }
                    }
                    total_tstep = tstep;
                    compute_tstep = 0;
                }
                tstep = total_tstep / rhof;
            }
            gegs_float2 dd = dedx[lelke - 1];
            float dedx0 = dd.x + dd.y * elke;
            dedx_val = rhof * dedx0;
            gegs_float2 tm = mt->tmxs[medium * MXEKE + lelke - 1];
            float tmxs_val = tm.x + tm.y * elke;
            tmxs_val /= rhof;
            float ekei = mt->e_array[medium * MXEKE + lelke - 1];
            float elkei = ((float)lelke - ek.x) / ek.y;
            range = drange(dedx, eke, elke, ekei, elkei, lelke);
            range = (range + mt->range_ep[(medium * MXEKE + lelke - 1) * 2 + qel]) / rhof;
            tmxs_val = fminf(tmxs_val, mt->smaxir[medium]);
            tustep = fminf(tstep, fminf(tmxs_val, range));
            tperp = hownear(c, p);
            p->dnear = tperp;
            /* total energy against e_max_rr (Q7); kinetic with GEGS_FIX_Q7 */
            const float e_rr = (c->fixes & GEGS_FIX_Q7) ? eke : p->e;
            if (mt->i_do_rr[medium] == 1 && e_rr < mt->e_max_rr[medium]) {
                if (tperp >= range) {
                    edep1 = (p->e - RM);
                    escore(c, p->region, edep1 * p->wt);
                    if (lelec > 0) annih_at_rest(c, s, p, &rd);
                    return;
                }
            }
            blccl = rhof * mt->blcc[medium];
            xccl = rhof * mt->xcc[medium];
            float p2 = eke * (eke + RMT2);
            beta2 = p2 / (p2 + RMSQ);
            float ssmfp = beta2 / blccl;
            float skindepth = mt->skindepth_for_bca * ssmfp;
            tustep = fminf(tustep, fmaxf(tperp, skindepth));
            if (tustep <= tperp && (!mt->exact_bca || tustep > skindepth)) {
                callhowfar = 0;
                domultiple = 0;
                dosingle = 0;
                callmsdist = 1;
                de = compute_eloss(mt, qel, medium, range, rhof, tustep, eke, elke, lelke);
                tvstep = tustep;
                /* transport_algorithm: 0 PRESTA-II, 1 PRESTA-I (the reference's default) */
                if (mt->transport_algorithm == 0) ustep = msdist_pII(mt, s, de, tustep, p, &rd);
                else ustep = msdist_pI(mt, s, de, tustep, p, &rd);
            } else if (mt->exact_bca) {
                /* exact BCA (electron.cu:1718-1752): cross the boundary in
                 * single-scattering mode; the distance to the next elastic
                 * scattering, in mean free paths, shortens the step */
                callmsdist = 0;
                domultiple = 0;
                float rnnoss = rnd(s);
                float lambda = rnnoss < 1.0f ? -om_logf(1.0f - rnnoss) : INFINITY;
                float tmp = eke / RM + 1.0f;
                float lambda_max = 0.5f * blccl * RM / dedx_val * tmp * tmp * tmp;
                if (lambda >= 0.0f && lambda_max > 0.0f) {
                    float tuss;
                    if (lambda < lambda_max) tuss = lambda * ssmfp * (1.0f - 0.5f * lambda / lambda_max);
                    else tuss = 0.5f * lambda * ssmfp;
                    if (tuss < tustep) {
                        tustep = tuss;
                        dosingle = 1;
                    } else {
                        dosingle = 0;
                    }
                } else {
                    /* the reference drops the electron (electron.cu:1742-1748) */
                    c->dropped += ((double)p->e + (double)p->charge * PRM) * (double)p->wt;
                    return;
                }
                ustep = tustep;
                callhowfar = ustep < tperp ? 0 : 1;
            } else {
                callmsdist = 0;
                /* inexact BCA (exact_bca = 0 in the reference defaults) */
                dosingle = 0;
                domultiple = 1;
                ekems = eke - 0.5f * tustep * dedx_val;
                float p2m = ekems * (ekems + RMT2);
                beta2 = p2m / (p2m + RMSQ);
                chia2 = xccl / (4.0f * blccl * p2m);
                xi = 0.5f * xccl / p2m / beta2 * tustep;
                xi = xi * (om_logf(1.0f + 1.0f / chia2) - 1.0f / (1.0f + chia2));
                if (xi < 0.1f) ustep = tustep * (1.0f - xi * (0.5f - xi * 0.166667f));
                else ustep = tustep * (1.0f - om_expf(-xi)) / xi;
                callhowfar = ustep < tperp ? 0 : 1;
            }
        }
        uint32_t irnew = p->region;
        int idisc = 0;
        float ustep0 = ustep;
        if (callhowfar || p->wt <= 0.0f) irnew = howfar(c, p, &ustep);
        if (irnew == 0) idisc = (rd.med == GEGS_VACUUM || p->leaf == PASSMLC) ? 1 : -1;
        if (idisc > 0) {
            c->escaped += ((double)p->e + (double)p->charge * PRM) * (double)p->wt;
            p->region = 0;
            return;
        }
        if (ustep <= 0.0f) ustep = 0.0f;
        if (ustep == 0.0f || medium == GEGS_VACUUM) {
            if (ustep != 0.0f) {
                vstep = ustep;
                tvstep = vstep;
                p->x = p->x + p->u * vstep;
                p->y = p->y + p->v * vstep;
                p->z = p->z + p->w * vstep;
                p->dnear = p->dnear - vstep;
            }
            if (p->region != irnew) c->cnt.region_gathers++;
            p->region = irnew;
            irl = irnew;
            rd = c->reg[irl];
            if (eie <= rd.ecut) {
                edep1 = (p->e - RM);
                escore(c, p->region, edep1 * p->wt);
                if (p->charge > 0) annih_at_rest(c, s, p, &rd);
                return;
            }
            if (ustep != 0.0f && idisc < 0) {
                edep1 = (p->e + (float)p->charge * RM);
                escore(c, p->region, edep1 * p->wt);
                return;
            }
            is_tstep = 1;
            continue;
        }
        vstep = ustep;
        if (callhowfar && mt->exact_bca) {
            /* single-scattering mode: a crossed boundary ends the step there
             * and shuts off the single scattering */
            tvstep = vstep;
            if (tvstep != tustep) dosingle = 0;
            de = compute_eloss(mt, qel, medium, range, rhof, tvstep, eke, elke, lelke);
        } else if (callhowfar) {
            if (vstep < ustep0) {
                ekems = eke - 0.5f * tustep * vstep / ustep0 * dedx_val;
                float p2m = ekems * (ekems + RMT2);
                beta2 = p2m / (p2m + RMSQ);
                chia2 = xccl / (4.0f * blccl * p2m);
                xi = 0.5f * xccl / p2m / beta2 * vstep;
                xi = xi * (om_logf(1.0f + 1.0f / chia2) - 1.0f / (1.0f + chia2));
                if (xi < 0.1f) {
                    tvstep = vstep * (1.0f + xi * (0.5f + xi * 0.333333f));
                } else {
                    if (xi < 0.999999f) tvstep = -vstep * om_logf(1.0f - xi) / xi;
                    else {         /* electron.cu:2031-2035: particle dropped */
                        c->dropped += ((double)p->e + (double)p->charge * PRM) * (double)p->wt;
                        return;
                    }
                }
            } else {
                tvstep = tustep;
            }
            de = compute_eloss(mt, qel, medium, range, rhof, tvstep, eke, elke, lelke);
        } else {
            tvstep = tustep;
            if (!callmsdist) de = compute_eloss(mt, qel, medium, range, rhof, tvstep, eke, elke, lelke);
        }
        float save_de = de;
        c->path += (double)tvstep;
        float edep = de;
        float ekef = eke - de;
        if (!callmsdist) {
            float sinthe = 0.0f, costhe = 1.0f;
            if (domultiple) {
                float lambda = blccl * tvstep / beta2 / (1.0f + chia2);
                mscat(mt, s, lambda, chia2, xi, &costhe, &sinthe);
            } else if (dosingle) {
                /* single scattering at the end of the step (electron.cu:2247-2276) */
                ekems = fmaxf(ekef, rd.ecut - RM);
                float p2s = ekems * (ekems + RMT2);
                chia2 = mt->xcc[medium] / (4.0f * mt->blcc[medium] * p2s);
                sscat(s, chia2, &costhe, &sinthe);
            }
            p->x = p->x + p->u * vstep;
            p->y = p->y + p->v * vstep;
            p->z = p->z + p->w * vstep;
            if (domultiple || dosingle) {
                float sinphi, cosphi;
                uphi21(s, costhe, sinthe, &cosphi, &sinphi, p);
            }
        }
        (void)ekef;
        p->dnear = p->dnear - vstep;
        irold = p->region;
        eie = eie - edep;
        p->e = eie;
        escore(c, irold, edep * p->wt);
        if (irnew == irl && eie <= rd.ecut) {
            edep1 = (p->e - RM);
            escore(c, p->region, edep1 * p->wt);
            if (p->charge > 0) annih_at_rest(c, s, p, &rd);
            return;
        }
        int medold = medium;
        if (medium != GEGS_VACUUM) {
            eke = eie - RM;
            elke = om_logf(eke);
            gegs_float2 ek = mt->eke01[medium];
            lelke = (int)(ek.x + ek.y * elke);
        }
        if (irnew != irold) {
            p->region = irnew;
            irl = irnew;
            rd = c->reg[irl];
            medium = rd.med;
            c->cnt.region_gathers++;
        }
        if (eie <= rd.ecut) {
            edep1 = (p->e - RM);
            escore(c, p->region, edep1 * p->wt);
            if (p->charge > 0) annih_at_rest(c, s, p, &rd);
            return;
        }
        if (idisc < 0) {
            edep1 = (p->e + (float)p->charge * RM);
            escore(c, p->region, edep1 * p->wt);
            return;
        }
        if (medium != medold) { is_tstep = 1; continue; }
        demfp = demfp - save_de * sig_val;
        total_tstep = total_tstep - tvstep * rhof;
        if (total_tstep < 1e-9f) demfp = 0.0f;
        if ((double)demfp >= EPSEMFP) { is_tstep = 0; continue; }
        gegs_float2 sd = mt->sig[qel * MXEKE * MXMED + medium * MXEKE + lelke - 1];
        gegs_float2 dd = mt->dedx[qel * MXEKE * MXMED + medium * MXEKE + lelke - 1];
        float sigf = sd.x + sd.y * elke;
        float dedx0 = dd.x + dd.y * elke;
        sigf /= dedx0;
        float sigratio = sigf / sig0;
        float rfict = rnd(s);
        if (rfict > sigratio) { is_tstep = 1; continue; }
        if (lelec < 0) {
            gegs_float2 b1 = mt->ebr1[medium * MXEKE + lelke - 1];
            float ebr1 = b1.x + b1.y * elke;
            float rnno24 = rnd(s);
            if (rnno24 <= ebr1) { push_cont(c, s, p, E_BREMS); return; }
            if (p->e <= mt->thmoll[medium] && mt->eii_flag == 0) {
                if (ebr1 <= 0.0f) { is_tstep = 1; continue; }
                push_cont(c, s, p, E_BREMS);
                return;
            }
            push_cont(c, s, p, E_MOLLER);
            return;
        }
        gegs_float2 pb1 = mt->pbr1[medium * MXEKE + lelke - 1];
        float pbr1 = pb1.x + pb1.y * elke;
        float rnno25 = rnd(s);
        if (rnno25 < pbr1) { push_cont(c, s, p, E_BREMS); return; }
        gegs_float2 pb2 = mt->pbr2[medium * MXEKE + lelke - 1];
        float pbr2 = pb2.x + pb2.y * elke;
        if (rnno25 < pbr2) push_cont(c, s, p, E_BHABHA);
        else push_cont(c, s, p, E_ANNIH);
        return;
    }
}

/* ------------------------------------------------------------ eii_sample (eii_sample.cu:4-96) */
/* ionisation of shell ish (binding uj) of element iz by the electron p: the
 * energy transfer from a mixture of the distant (Bethe) and close (Moller)
 * collision shapes, the secondary created above AE, then relaxation of the
 * vacancy; *edep is set (not accumulated), as the reference does */
static void eii_sample(octx *c, ostream *s, int ish, int iz, float uj, opart *p, const gegs_region *rd, float *edep) {
    const gegs_media_tables *mt = c->mt;
    int medium = rd->med;
    opart p1 = *p;
    float eie = p->e;
    float t = eie - RM;
    float tau = t / RM;
    float tau12 = (tau + 1) * (tau + 1);
    float p2 = tau * tau + 2 * tau;
    float beta2 = p2 / tau12;
    float wmax = (float)(0.5 * (double)(t + uj));
    float xmax = uj / wmax;
    float c1 = (wmax / eie) * (wmax / eie);
    float c2 = (2 * tau + 1) / tau12;
    float fm_s = om_logf(RMT2 * p2 / uj) - beta2 - 0.5f;
    float prob_s = (float)(0.66666667 * (double)fm_s * (double)(1 + xmax + xmax * xmax));
    float fm_h = 2 + c1 - c2;
    if (fm_h < 1) fm_h = 1;
    float prob = fm_h + prob_s;
    float r1, frej, wx;
    do {
        r1 = rnd(s);
        if ((r1 * prob) < fm_h) {
            r1 = rnd(s);
            wx = 1 / (r1 * xmax + 1 - r1);
            float wxx = wx * xmax;
            float aux = wxx / (2 - wxx);
            frej = (1 + aux * (aux - c2) + c1 * wxx * wxx) / fm_h;
        } else {
            r1 = rnd(s);
            /* powf(x, 0.333333333f) through the library's own log and exp */
            wx = 1.0f / om_expf(om_logf(r1 * xmax * xmax * xmax + 1 - r1) * 0.333333333f);
            frej = 1 - om_logf(wx) / fm_s;
        }
        r1 = rnd(s);
    } while (r1 >= frej);
    wx = wx * uj;
    float h1 = (eie + RM) / t;
    float ese1 = eie - wx;
    p->e = ese1;
    float dcosth = h1 * (ese1 - RM) / (ese1 + RM);
    float sinthe = sqrtf(1 - dcosth);
    float costhe = sqrtf(dcosth);
    float sinphi, cosphi;
    uphi21(s, costhe, sinthe, &cosphi, &sinphi, p);
    float ese2 = wx - uj + RM;
    if (ese2 > mt->ae[medium]) {
        p1.e = ese2;
        dcosth = h1 * (ese2 - RM) / (ese2 + RM);
        sinthe = -sqrtf(1 - dcosth);
        costhe = sqrtf(dcosth);
        uphi32(costhe, sinthe, cosphi, sinphi, &p1);
        push_child(c, s, &p1, E_ELECTRON_STEP);
        *edep = 0;
    } else {
        *edep = wx - uj;
    }
    relax(c, s, uj, ish, iz, edep, *p, rd);
}

/* is the Moller interaction an ionisation of an EII shell? (moller.cu:78-121)
 * 1 with the shell, its element and binding energy, or 0 (one uniform drawn
 * whenever the medium has EII shells) */
static int eii_choose(const gegs_media_tables *mt, ostream *s, int medium, float ekin, int *ish_out, int *iz_out,
                      float *uj_out) {
    if (!(mt->eii_flag > 0 && mt->eii_nsh[medium] > 0)) return 0;
    float elke = om_logf(ekin);
    unsigned lelke = (unsigned)(mt->eke01[medium].y * elke + mt->eke01[medium].x);
    gegs_float2 sg = mt->sig[medium * MXEKE + lelke - 1];
    gegs_float2 eb = mt->ebr1[medium * MXEKE + lelke - 1];
    float sigm = sg.y * elke + sg.x;
    float pbrem = eb.y * elke + eb.x;
    sigm = sigm * (1 - pbrem);
    float rsh = rnd(s);
    rsh = sigm * rsh;
    for (int iele = 0; iele < mt->nne[medium]; iele++) {
        int iz = (int)((double)mt->zelem[iele * MXMED + medium] + 0.5);
        int nsh = mt->eii_no[iele * MXMED + medium];
        if (nsh <= 0) continue;
        int ifirst = mt->eii_first[iele * MXMED + medium];
        for (int ish = 1; ish <= nsh; ish++) {
            float uj = mt->binding_energies[(iz - 1) * MXSHELL + ish - 1];
            if ((ekin > uj) && (uj > mt->te[medium] || uj > mt->ap[medium])) {
                int jj = ifirst + ish - 1;
                int i = (int)(mt->eii[jj - 1].x * elke + mt->eii[jj - 1].y + (float)((jj - 1) * N_EII_BINS));
                float sig_j = mt->eii_xsection[i - 1].x * elke + mt->eii_xsection[i - 1].y;
                sig_j = sig_j * mt->pz[MXMED * iele + medium] * mt->eii_cons[medium];
                rsh -= sig_j;
                if (rsh < 0.0f) {
                    *ish_out = ish;
                    *iz_out = iz;
                    *uj_out = uj;
                    return 1;
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------ moller (moller.cu:16-210) */
static void moller(octx *c, ostream *s, opart *p) {
    const gegs_media_tables *mt = c->mt;
    gegs_region rd = c->reg[p->region];
    int medium = rd.med;
    float edep = 0.0f;
    opart p1 = *p;
    float eie = p->e;
    float ekin = eie - RM;
    int ish, iz;
    float uj;
    if (eii_choose(mt, s, medium, ekin, &ish, &iz, &uj)) {
        eii_sample(c, s, ish, iz, uj, p, &rd, &edep);
        if (p->e < rd.ecut) edep += p->e - RM;
        else push_child(c, s, p, E_ELECTRON_STEP);
        escore(c, p->region, edep * p->wt);
        return;
    }
    if (ekin <= 2.0f * mt->te[medium]) {
        edep = (p->e - RM);
        escore(c, p->region, edep * p->wt);
        return;
    }
    float t0 = ekin / RM;
    float e0 = t0 + 1.0f;
    float extrae = eie - mt->thmoll[medium];
    float e02 = e0 * e0;
    float g2 = t0 * t0 / e02;
    float g3 = (2.0f * t0 + 1.0f) / e02;
    float gmax = 1.0f + 1.25f * g2;
    float rnno27, rejf4, br, r;
    do {
        rnno27 = rnd(s);
        br = mt->te[medium] / (ekin - extrae * rnno27);
        r = br / (1.0f - br);
        rnno27 = rnd(s);
        rejf4 = 1.0f + g2 * br * br + r * (r - g3);
        rnno27 = gmax * rnno27;
    } while (rnno27 > rejf4);
    float ekse2 = br * ekin;
    float ese1 = eie - ekse2;
    float ese2 = ekse2 + RM;
    p->e = ese1;
    p1.e = ese2;
    float h1 = (eie + RM) / ekin;
    float dcosth = h1 * (ese1 - RM) / (ese1 + RM);
    float sinthe = sqrtf(1.0f - dcosth);
    float costhe = sqrtf(dcosth);
    float sinphi, cosphi;
    uphi21(s, costhe, sinthe, &cosphi, &sinphi, p);
    if (p->e < rd.ecut) edep += (p->e - RM);
    else push_child(c, s, p, E_ELECTRON_STEP);
    dcosth = h1 * (ese2 - RM) / (ese2 + RM);
    sinthe = -sqrtf(1.0f - dcosth);
    costhe = sqrtf(dcosth);
    uphi32(costhe, sinthe, cosphi, sinphi, &p1);
    if (p1.e < rd.ecut) edep += (p1.e - RM);
    else push_child(c, s, &p1, E_ELECTRON_STEP);
    escore(c, p->region, edep * p->wt);
}

/* ------------------------------------------------------------ bhabha (bhabha.cu:12-172) */
static void bhabha(octx *c, ostream *s, opart *p) {
    const gegs_media_tables *mt = c->mt;
    gegs_region rd = c->reg[p->region];
    int medium = rd.med;
    float edep = 0.0f;
    opart p1 = *p;
    float eip = p->e;
    float ekin = eip - RM;
    float t0 = ekin / RM;
    float e0 = t0 + 1.0f;
    float yy = 1.0f / (t0 + 2.0f);
    float e02 = e0 * e0;
    float beta2 = (e02 - 1.0f) / e02;
    float ep0 = mt->te[medium] / ekin;
    float ep0c = 1.0f - ep0;
    float y2 = yy * yy;
    float yp = 1.0f - 2.0f * yy;
    float yp2 = yp * yp;
    float b4 = yp2 * yp;
    float b3 = b4 + yp2;
    float b2 = yp * (3.0f + y2);
    float b1 = 2.0f - y2;
    float rnno03, rejf2, br;
    do {
        rnno03 = rnd(s);
        br = ep0 / (1.0f - ep0c * rnno03);
        rnno03 = rnd(s);
        rejf2 = 1.0f - beta2 * br * (b1 - br * (b2 - br * (b3 - br * b4)));
    } while (rnno03 > rejf2);
    if (br < 0.5f) {
        p1.charge = -1;
    } else {
        p->charge = -1;
        p1.charge = 1;
        br = 1.0f - br;
    }
    br = fmaxf(br, 0.0f);
    float ekse2 = br * ekin;
    float ese1 = eip - ekse2;
    float ese2 = ekse2 + RM;
    p->e = ese1;
    p1.e = ese2;
    float h1 = (eip + RM) / ekin;
    float dcosth = fminf(1.0f, h1 * (ese1 - RM) / (ese1 + RM));
    float sinthe = sqrtf(1.0f - dcosth);
    float costhe = sqrtf(dcosth);
    float sinphi, cosphi;
    uphi21(s, costhe, sinthe, &cosphi, &sinphi, p);
    if (p->e < rd.ecut) {
        edep += (p->e - RM);
        if (p->charge > 0) annih_at_rest(c, s, p, &rd);
    } else {
        push_child(c, s, p, E_ELECTRON_STEP);
    }
    dcosth = h1 * (ese2 - RM) / (ese2 + RM);
    sinthe = -sqrtf(1.0f - dcosth);
    costhe = sqrtf(dcosth);
    uphi32(costhe, sinthe, cosphi, sinphi, &p1);
    if (p1.e < rd.ecut) {
        edep += (p1.e - RM);
        if (p1.charge > 0) annih_at_rest(c, s, &p1, &rd);
    } else {
        push_child(c, s, &p1, E_ELECTRON_STEP);
    }
    escore(c, p->region, edep * p->wt);
}

/* ------------------------------------------------------------ brems (brems.cu:36-308), ibrdst = 1 (KM) */
static void brems(octx *c, ostream *s, opart *p) {
    const gegs_media_tables *mt = c->mt;
    gegs_region rd = c->reg[p->region];
    int medium = rd.med;
    float edep = 0.0f;
    opart p1 = *p;
    if (mt->nbr_split < 1) return;
    float eie = p->e;
    int l = eie < 50.0f ? 1 : 3;
    int l1 = l + 1;
    float ekin = eie - RM;
    float elke = om_logf(ekin);
    float brmin = mt->ap[medium] / ekin;
    float waux = elke - mt->log_ap[medium];
    float a = p->u, b = p->v, cc = p->w;
    float sinpsi = a * a + b * b, sindel = 0.0f, cosdel = 0.0f;
    if (sinpsi > 1e-20f) {
        sinpsi = sqrtf(sinpsi);
        sindel = b / sinpsi;
        cosdel = a / sinpsi;
    }
    float ztarg = mt->zbrang[medium];
    float tteie = eie / RM;
    float beta = sqrtf((tteie - 1.0f) * (tteie + 1.0f)) / tteie;
    float y2max = 2.0f * beta * (1.0f + beta) * tteie * tteie;
    float y2maxi = 1.0f / y2max;
    float z2max = y2max + 1.0f;
    float z2maxi = sqrtf(z2max);
    float ese = eie;
    for (int ibr = 1; ibr <= mt->nbr_split; ++ibr) {
        float rnno06, rejf, esg;
        do {
            rnno06 = rnd(s);
            float br = brmin * om_expf(rnno06 * waux);
            esg = ekin * br;
            ese = eie - esg;
            float delta = esg / eie / ese * mt->delcm[medium];
            float aux = ese / eie;
            float phi1, phi2;
            if (delta < 1.0f) {
                phi1 = mt->dl1[medium * 8 + l - 1] + delta * (mt->dl2[medium * 8 + l - 1] + delta * mt->dl3[medium * 8 + l - 1]);
                phi2 = mt->dl1[medium * 8 + l1 - 1] + delta * (mt->dl2[medium * 8 + l1 - 1] + delta * mt->dl3[medium * 8 + l1 - 1]);
            } else {
                phi1 = mt->dl4[medium * 8 + l - 1] + mt->dl5[medium * 8 + l - 1] * om_logf(delta + mt->dl6[medium * 8 + l - 1]);
                phi2 = phi1;
            }
            rejf = (1.0f + aux * aux) * phi1 - 2.0f * aux * phi2 / 3.0f;
            rnno06 = rnd(s);
        } while (rnno06 >= rejf);
        p1.e = esg;
        p1.charge = 0;
        p1.wt = p->wt / (float)mt->nbr_split;
        /* Koch-Motz 2BS angle (brems.cu:191-233) */
        float y2tst;
        {
            float esedei = ese / RM / tteie;
            float rjarg1 = 1.0f + esedei * esedei;
            float rjarg2 = rjarg1 + 2.0f * esedei;
            float aux = 2.0f * ese * tteie / esg;
            aux = aux * aux;
            float aux1 = aux * ztarg;
            float rjarg3;
            if (aux1 > 10.0f) rjarg3 = mt->lzbrang[medium] + (1.0f - aux1) / (aux1 * aux1);
            else rjarg3 = om_logf(aux / (1.0f + aux1));
            float rejmax = rjarg1 * rjarg3 - rjarg2;
            float rtest, rejtst;
            do {
                y2tst = rnd(s);
                rtest = rnd(s);
                float aux3 = z2maxi / (y2tst + (1.0f - y2tst) * z2maxi);
                rtest = rtest * aux3 * rejmax;
                y2tst = aux3 * aux3 - 1.0f;
                float y2tst1 = esedei * y2tst / (aux3 * aux3 * aux3 * aux3);
                float aux4 = 16.0f * y2tst1 - rjarg2;
                float aux5 = rjarg1 - 4.0f * y2tst1;
                if (rtest < aux4 + aux5 * rjarg3) break;
                float aux2 = om_logf(aux / (1.0f + aux1 / (aux3 * aux3 * aux3 * aux3)));
                rejtst = aux4 + aux5 * aux2;
            } while (rtest >= rejtst);
        }
        float costhe = 1.0f - 2.0f * y2tst * y2maxi;
        float sinthe = sqrtf(fmaxf((1.0f - costhe) * (1.0f + costhe), 0.0f));
        float cphi, sphi;
        select_azimuth(s, &cphi, &sphi);
        if (sinpsi >= 1e-10f) {
            float us = sinthe * cphi;
            float vs = sinthe * sphi;
            p1.u = cc * cosdel * us - sindel * vs + a * costhe;
            p1.v = cc * sindel * us + cosdel * vs + b * costhe;
            p1.w = cc * costhe - sinpsi * us;
        } else {
            p1.u = sinthe * cphi;
            p1.v = sinthe * sphi;
            p1.w = cc * costhe;
        }
        if (p1.e < rd.pcut) edep += p1.e;
        else push_child(c, s, &p1, P_PHOTON_STEP);
    }
    p->e = ese;
    if (p->e < rd.ecut) edep += (p->e - RM);
    else push_child(c, s, p, E_ELECTRON_STEP);
    escore(c, p->region, edep * p->wt);
}

/* ------------------------------------------------------------ annih (annih.cu:16-206) */
static void annih(octx *c, ostream *s, opart *p) {
    const gegs_media_tables *mt = c->mt;
    gegs_region rd = c->reg[p->region];
    float edep = 0.0f;
    if (mt->nbr_split <= 0) return;      /* Q5 */
    float avip = p->e + RM;
    float a = avip / RM;
    float g = a - 1.0f;
    float t = g - 1.0f;
    float pp = sqrtf(a * t);
    float pot = pp / t;
    float ep0 = 1.0f / (a + pp);
    float wsamp = om_logf((1.0f - ep0) / ep0);
    float aa = p->u, bb = p->v, cc = p->w;
    float sinpsi = aa * aa + bb * bb, sindel = 0.0f, cosdel = 0.0f;
    if (sinpsi > 1e-20f) {
        sinpsi = sqrtf(sinpsi);
        sindel = bb / sinpsi;
        cosdel = aa / sinpsi;
    }
    if (mt->nbr_split > 1) p->wt = p->wt / (float)mt->nbr_split;
    for (int ibr = 1; ibr <= mt->nbr_split; ++ibr) {
        float rnno01, rejf, ep;
        do {
            rnno01 = rnd(s);
            ep = ep0 * om_expf(rnno01 * wsamp);
            rnno01 = rnd(s);
            rejf = 1.0f - (ep * a - 1.0f) * (ep * a - 1.0f) / (ep * (a * a - 2.0f));
        } while (rnno01 > rejf);
        float esg1 = avip * ep;
        p->e = esg1;
        p->charge = 0;
        float costhe = fminf(1.0f, (esg1 - RM) * pot / esg1);
        float sinthe = sqrtf(1.0f - costhe * costhe);
        float cphi, sphi;
        select_azimuth(s, &cphi, &sphi);
        if (sinpsi >= 1e-10f) {
            float us = sinthe * cphi, vs = sinthe * sphi;
            p->u = cc * cosdel * us - sindel * vs + aa * costhe;
            p->v = cc * sindel * us + cosdel * vs + bb * costhe;
            p->w = cc * costhe - sinpsi * us;
        } else {
            p->u = sinthe * cphi;
            p->v = sinthe * sphi;
            p->w = cc * costhe;
        }
        if (p->e < rd.pcut) edep += p->e;
        else push_child(c, s, p, P_PHOTON_STEP);
        float esg2 = avip - esg1;
        p->e = esg2;
        costhe = fminf(1.0f, (esg2 - RM) * pot / esg2);
        sinthe = -sqrtf(1.0f - costhe * costhe);
        if (sinpsi >= 1e-10f) {
            float us = sinthe * cphi, vs = sinthe * sphi;
            p->u = cc * cosdel * us - sindel * vs + aa * costhe;
            p->v = cc * sindel * us + cosdel * vs + bb * costhe;
            p->w = cc * costhe - sinpsi * us;
        } else {
            p->u = sinthe * cphi;
            p->v = sinthe * sphi;
            p->w = cc * costhe;
        }
        if (p->e < rd.pcut) edep += p->e;
        else push_child(c, s, p, P_PHOTON_STEP);
    }
    escore(c, p->region, edep * p->wt);
}

/* ------------------------------------------------------------ driver */
uint64_t oracle_primary_lineage(uint64_t hist) { return mix64(hist ^ 0x5851F42D4C957F2Dull); }

size_t oracle_sizeof_media_tables(void) { return sizeof(gegs_media_tables); }

/* Transport histories [i0, i1) of the source with one context (one thread). */
static void run_range(octx *c, const gegs_source *src, uint64_t i0, uint64_t i1, uint64_t hist_offset,
                      double *hist_edep, uint32_t *hist_steps) {
    for (uint64_t i = i0; i < i1; ++i) {
        c->hist = hist_offset + i;
        c->cur_edep = 0.0;
        c->cur_steps = 0;
        c->cnt.histories++;
        opart p;
        memset(&p, 0, sizeof(p));
        p.e = src[i].e; p.x = src[i].x; p.y = src[i].y; p.z = src[i].z;
        p.u = src[i].u; p.v = src[i].v; p.w = src[i].w; p.wt = src[i].wt;
        p.charge = src[i].charge;
        p.region = 0;
        p.module = c->f.module_first;
        p.leaf = OUTMLC;
        p.bank = -1;
        p.dnear = INFI_DIST;
        p.lin = oracle_primary_lineage(c->hist);
        p.blk = 0;
        p.status = p.charge == 0 ? P_PHOTON_STEP : E_ELECTRON_STEP;
        c->sp = 0;
        push(c, &p);
        while (c->sp > 0) {
            opart q = c->stack[--c->sp];
            ostream s;
            st_begin(&s, &q, c->seed);
            c->cnt.interactions[q.status]++;
            switch (q.status) {
            case P_PHOTON_STEP: photon_step(c, &s, &q); break;
            case P_COMPTON: compton(c, &s, &q); break;
            case P_PHOTO: photo(c, &s, &q); break;
            case P_RAYLEIGH: rayleigh(c, &s, &q); break;
            case P_PAIR: pair(c, &s, &q); break;
            case E_ELECTRON_STEP: electron_step(c, &s, &q); break;
            case E_MOLLER: moller(c, &s, &q); break;
            case E_BREMS: brems(c, &s, &q); break;
            case E_BHABHA: bhabha(c, &s, &q); break;
            case E_ANNIH: annih(c, &s, &q); break;
            default: break;
            }
        }
        if (hist_edep) hist_edep[i] = c->cur_edep;
        if (hist_steps) hist_steps[i] = c->cur_steps;
    }
    size_t nvox = (size_t)c->nx * c->ny * c->nz;
    for (size_t v = 0; v < nvox; ++v)
        if (c->last[v]) { c->dose2[v] += c->tmp[v] * c->tmp[v]; c->last[v] = 0; c->tmp[v] = 0.0; }
}

/*
 * Transport n source particles, one history each (history index
 * hist_offset + i).  dose: FP64 energy per voxel; dose2: history-by-history
 * sum of squared per-history voxel energies.  hist_edep / hist_steps
 * (optional): per-history energy scored in the phantom and step count.
 * balance (optional): [escaped, dropped] energy.  Bounds arrays carry
 * nx+2, ny+2, nz+2 entries (see Q1 above).  nthreads > 1 splits the
 * histories into contiguous ranges, one per OpenMP thread, each with its
 * own tallies, summed at the end; per-history results do not depend on it.
 */
/* leaf ends of a field (init_leaf_pos, init.cu:1594-1662): an open pair's
 * bank A end moves to -pos[j+60], bank B's to pos[j], scaled from the
 * isocentre plane to the leaf centre plane */
void oracle_leaf_ends(const gegs_mlc *m, const float *pos, float sid, gegs_float2 *out) {
    float zc = m->zmin + m->zthick / 2.0f;
    for (int j = 0; j < NLEAF; ++j) {
        gegs_float2 *en = out + j * 34;
        for (int i = 0; i < 34; ++i) en[i] = m->cross_yz[i];
        if (pos[j] + pos[j + NLEAF] > 0.0f) {
            float a = -pos[j + NLEAF], b = pos[j];
            for (int i = 2; i < 17; ++i) en[i].x = m->cross_yz[i].x + a * zc / sid;
            for (int i = 17; i < 32; ++i) en[i].x = m->cross_yz[i].x + b * zc / sid;
        }
    }
}

/* MLC probes for the geometry tests */
int oracle_leaf_index(const gegs_mlc *m, float x, float z) { return leaf_index(m, mk2(x, z)); }
uint32_t oracle_howfar_mlc(const gegs_mlc *m, const gegs_float2 *ends, const gegs_particle *in, int32_t *leaf,
                           int32_t *bank, int32_t *module, float *t) {
    octx c;
    memset(&c, 0, sizeof(c));
    c.mlc = m;
    c.endnodes = ends;
    set_module_order(&c, m, NULL);
    opart p;
    memset(&p, 0, sizeof(p));
    p.x = in->x; p.y = in->y; p.z = in->z; p.u = in->u; p.v = in->v; p.w = in->w;
    p.region = in->region; p.leaf = *leaf; p.bank = *bank; p.module = GEGS_M_MLC;
    uint32_t r = howfar_mlc(&c, &p, t);
    *leaf = p.leaf; *bank = p.bank; *module = p.module;
    return r;
}
/* jaw polygons of a field (setup_secjaws, init.cu:1504-1592); the +-0.5 cm
 * margins are double literals in the reference */
void oracle_jaw_polys(const gegs_jaws *g, const float *xy, float sid, gegs_float2 *out) {
    float zmin[2] = {g->zminy, g->zminx}, zmax[2] = {g->zmaxy, g->zmaxx};
    for (int k = 0; k < 2; ++k) {
        float a1 = xy[0] * zmin[k] / sid, a2 = xy[1] * zmin[k] / sid;
        float c1 = xy[2] * zmin[k] / sid, c2 = xy[3] * zmin[k] / sid;
        gegs_float2 *bm = out + 4 * k, *bx = out + 8 + 4 * k, *e = out + 16 + 8 * k;
        bm[0].x = (float)(a1 - 0.5); bm[0].y = (float)(c1 - 0.5);
        bm[1].x = (float)(a2 + 0.5); bm[1].y = (float)(c1 - 0.5);
        bm[2].x = (float)(a2 + 0.5); bm[2].y = (float)(c2 + 0.5);
        bm[3].x = (float)(a1 - 0.5); bm[3].y = (float)(c2 + 0.5);
        bx[0].x = -g->rmax; bx[0].y = zmin[k];
        bx[1].x = g->rmax;  bx[1].y = zmin[k];
        bx[2].x = g->rmax;  bx[2].y = zmax[k];
        bx[3].x = -g->rmax; bx[3].y = zmax[k];
        float lo = k == 0 ? xy[2] : xy[0], hi = k == 0 ? xy[3] : xy[1];
        e[0].x = -g->rmax; e[0].y = zmax[k];
        e[1].x = -g->rmax; e[1].y = zmin[k];
        e[2].x = lo * zmin[k] / sid; e[2].y = zmin[k];
        e[3].x = lo * zmax[k] / sid; e[3].y = zmax[k];
        e[4].x = hi * zmax[k] / sid; e[4].y = zmax[k];
        e[5].x = hi * zmin[k] / sid; e[5].y = zmin[k];
        e[6].x = g->rmax; e[6].y = zmin[k];
        e[7].x = g->rmax; e[7].y = zmax[k];
    }
}

/* jaw probe for the geometry tests: howfar of one particle in the Y (is_y)
 * or X pair; returns the region, updates t and module */
uint32_t oracle_howfar_jaws(const gegs_float2 *polys, int is_y, int jawx_next, const gegs_particle *in,
                            int32_t *module, float *t) {
    octx c;
    memset(&c, 0, sizeof(c));
    c.jaws = polys;
    c.jawx_next = jawx_next;
    c.fixes = g_fixes;
    opart p;
    memset(&p, 0, sizeof(p));
    p.x = in->x; p.y = in->y; p.z = in->z; p.u = in->u; p.v = in->v; p.w = in->w;
    p.region = in->region; p.module = is_y ? GEGS_M_SECJAWY : GEGS_M_SECJAWX;
    uint32_t r = howfar_jaws(&c, &p, t, is_y);
    *module = p.module;
    return r;
}

uint32_t oracle_where_mlc(const gegs_mlc *m, const gegs_float2 *ends, float x, float y, float z) {
    int index = OUTMLC, bank = -1;
    return mlc_where(m, ends, x, y, z, &index, &bank);
}

int oracle_run_mt(const gegs_media_tables *mt, const gegs_region *regions,
                  uint32_t nx, uint32_t ny, uint32_t nz,
                  const float *xb, const float *yb, const float *zb,
                  const gegs_field *field, const gegs_source *src, uint64_t n,
                  uint64_t hist_offset, uint64_t seed,
                  double *dose, double *dose2, double *hist_edep, uint32_t *hist_steps,
                  gegs_counters *counters, double *balance, int nthreads,
                  const gegs_mlc *mlc, const gegs_float2 *endnodes, const gegs_float2 *jaws,
                  const gegs_block_wedge *bw) {
    size_t nvox = (size_t)nx * ny * nz;
    if (nthreads < 1) nthreads = 1;
    if ((uint64_t)nthreads > n && n > 0) nthreads = (int)n;
    octx **cs = (octx **)calloc((size_t)nthreads, sizeof(octx *));
    int ok = 1;
    for (int t = 0; t < nthreads; ++t) {
        octx *c = (octx *)calloc(1, sizeof(octx));
        cs[t] = c;
        if (!c) { ok = 0; continue; }
        c->mt = mt; c->reg = regions; c->xb = xb; c->yb = yb; c->zb = zb;
        c->nx = nx; c->ny = ny; c->nz = nz; c->f = *field; c->seed = seed;
        c->mlc = mlc; c->endnodes = endnodes;
        c->jaws = jaws;
        set_module_order(c, mlc, bw);
        c->dose = t == 0 ? dose : (double *)calloc(nvox, sizeof(double));
        c->dose2 = t == 0 ? dose2 : (double *)calloc(nvox, sizeof(double));
        c->tmp = (double *)calloc(nvox, sizeof(double));
        c->last = (uint64_t *)calloc(nvox, sizeof(uint64_t));
        if (!c->dose || !c->dose2 || !c->tmp || !c->last) ok = 0;
    }
    if (ok) {
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
        for (int t = 0; t < nthreads; ++t) {
            uint64_t i0 = n * (uint64_t)t / (uint64_t)nthreads;
            uint64_t i1 = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
            run_range(cs[t], src, i0, i1, hist_offset, hist_edep, hist_steps);
        }
    }
    gegs_counters tot;
    memset(&tot, 0, sizeof(tot));
    double bal0 = 0.0, bal1 = 0.0;
    for (int t = 0; t < nthreads; ++t) {
        octx *c = cs[t];
        if (!c) continue;
        if (ok && t > 0) {
            for (size_t v = 0; v < nvox; ++v) { dose[v] += c->dose[v]; dose2[v] += c->dose2[v]; }
        }
        tot.photon_steps += c->cnt.photon_steps;
        tot.electron_steps += c->cnt.electron_steps;
        tot.region_gathers += c->cnt.region_gathers;
        tot.deposits += c->cnt.deposits;
        for (int k = 0; k < 10; ++k) tot.interactions[k] += c->cnt.interactions[k];
        tot.histories += c->cnt.histories;
        tot.overflow += c->cnt.overflow;
        bal0 += c->escaped;
        bal1 += c->dropped;
        if (t > 0) { free(c->dose); free(c->dose2); }
        free(c->tmp);
        free(c->last);
        free(c);
    }
    free(cs);
    if (!ok) return GEGS_ERR_ARG;
    if (counters) *counters = tot;
    if (balance) { balance[0] = bal0; balance[1] = bal1; }
    return GEGS_OK;
}

int oracle_run(const gegs_media_tables *mt, const gegs_region *regions,
               uint32_t nx, uint32_t ny, uint32_t nz,
               const float *xb, const float *yb, const float *zb,
               const gegs_field *field, const gegs_source *src, uint64_t n,
               uint64_t hist_offset, uint64_t seed,
               double *dose, double *dose2, double *hist_edep, uint32_t *hist_steps,
               gegs_counters *counters, double *balance) {
    return oracle_run_mt(mt, regions, nx, ny, nz, xb, yb, zb, field, src, n, hist_offset, seed,
                         dose, dose2, hist_edep, hist_steps, counters, balance, 1, NULL, NULL, NULL, NULL);
}

/* Single-sampler entry point for uniform-for-uniform checks: the handler of
 * in->status runs on one particle with the stream a transport run would give
 * it (lineage in->lin, block in->blk); the particles it pushes are returned
 * in push order, and *edep receives the energy it scored. */
static double g_last_path = 0.0;

int oracle_interaction(const gegs_media_tables *mt, const gegs_region *regions,
                       uint32_t nx, uint32_t ny, uint32_t nz,
                       const float *xb, const float *yb, const float *zb,
                       const gegs_field *field, const gegs_particle *in, uint64_t seed,
                       gegs_particle *out, int max_out, double *edep_out) {
    octx *c = (octx *)calloc(1, sizeof(octx));
    set_module_order(c, NULL, NULL);
    size_t nvox = (size_t)nx * ny * nz;
    double *dose = (double *)calloc(nvox, sizeof(double));
    double *dose2 = (double *)calloc(nvox, sizeof(double));
    c->mt = mt; c->reg = regions; c->xb = xb; c->yb = yb; c->zb = zb;
    c->nx = nx; c->ny = ny; c->nz = nz; c->f = *field; c->seed = seed;
    c->fixes = g_fixes;
    c->dose = dose; c->dose2 = dose2;
    c->tmp = (double *)calloc(nvox, sizeof(double));
    c->last = (uint64_t *)calloc(nvox, sizeof(uint64_t));
    opart p;
    memset(&p, 0, sizeof(p));
    p.e = in->e; p.x = in->x; p.y = in->y; p.z = in->z;
    p.u = in->u; p.v = in->v; p.w = in->w; p.wt = in->wt; p.dnear = in->dnear;
    p.charge = in->charge; p.region = in->region; p.module = field->module_first;
    p.leaf = OUTMLC; p.bank = -1;
    p.lin = in->lin; p.blk = in->blk; p.status = in->status;
    ostream s;
    st_begin(&s, &p, seed);
    switch (p.status) {
    case P_PHOTON_STEP: photon_step(c, &s, &p); break;
    case P_COMPTON: compton(c, &s, &p); break;
    case P_PHOTO: photo(c, &s, &p); break;
    case P_RAYLEIGH: rayleigh(c, &s, &p); break;
    case P_PAIR: pair(c, &s, &p); break;
    case E_ELECTRON_STEP: electron_step(c, &s, &p); break;
    case E_MOLLER: moller(c, &s, &p); break;
    case E_BREMS: brems(c, &s, &p); break;
    case E_BHABHA: bhabha(c, &s, &p); break;
    case E_ANNIH: annih(c, &s, &p); break;
    default: break;
    }
    int n = c->sp < max_out ? c->sp : max_out;
    for (int i = 0; i < n; ++i) {
        const opart *q = &c->stack[i];
        gegs_particle *o = out + i;
        memset(o, 0, sizeof(*o));
        o->e = q->e; o->x = q->x; o->y = q->y; o->z = q->z; o->u = q->u; o->v = q->v; o->w = q->w;
        o->wt = q->wt; o->dnear = q->dnear; o->region = q->region; o->charge = q->charge;
        o->status = q->status; o->lin = q->lin; o->blk = q->blk;
    }
    double tot = 0.0;
    for (size_t v = 0; v < nvox; ++v) tot += dose[v];
    *edep_out = tot;
    g_last_path = c->path;
    free(c->tmp); free(c->last); free(dose); free(dose2); free(c);
    return n;
}

/* electron path length (sum of the condensed-history steps' tvstep) of the
 * last oracle_interaction call; for the CSDA range checks (single-threaded use) */
double oracle_last_path(void) { return g_last_path; }

/* elementary functions, exported for their accuracy tests */
/* mscat probe: n draws of the multiple-scattering angle for fixed (lambda,
 * chia2, q1), stream i keyed by lineage lin0 + i; returns cos(theta) */
void oracle_mscat(const gegs_media_tables *mt, float lambda, float chia2, float q1, uint64_t seed, uint64_t lin0,
                  uint32_t n, float *cost) {
    for (uint32_t i = 0; i < n; ++i) {
        opart p;
        memset(&p, 0, sizeof(p));
        p.lin = lin0 + i;
        ostream s;
        st_begin(&s, &p, seed);
        float c, sn;
        mscat(mt, &s, lambda, chia2, q1, &c, &sn);
        cost[i] = c;
    }
}

/* msdist_pI probe: n draws of one PRESTA-I step (path tustep, energy loss
 * eloss) of an electron of total energy e starting at the origin with
 * direction (u, v, w) in a region of `medium` at relative density rhof, stream
 * i keyed by lineage lin0 + i; out[7 i ...] = displacement x, y, z, final
 * direction u, v, w and the returned straight-line distance ustep */
void oracle_msdist(const gegs_media_tables *mt, int medium, float rhof, float e, float eloss, float tustep,
                   const float *dir, uint64_t seed, uint64_t lin0, uint32_t n, float *out) {
    gegs_region rd;
    memset(&rd, 0, sizeof(rd));
    rd.med = (uint16_t)medium;
    rd.rhof = rhof;
    for (uint32_t i = 0; i < n; ++i) {
        opart p;
        memset(&p, 0, sizeof(p));
        p.e = e;
        p.u = dir[0]; p.v = dir[1]; p.w = dir[2];
        p.lin = lin0 + i;
        ostream s;
        st_begin(&s, &p, seed);
        /* the tables' transport algorithm: 0 PRESTA-II, 1 PRESTA-I */
        float us = mt->transport_algorithm == 0 ? msdist_pII(mt, &s, eloss, tustep, &p, &rd)
                                                : msdist_pI(mt, &s, eloss, tustep, &p, &rd);
        float *o = out + 7 * (size_t)i;
        o[0] = p.x; o[1] = p.y; o[2] = p.z; o[3] = p.u; o[4] = p.v; o[5] = p.w; o[6] = us;
    }
}

/* compute_eloss probe: energy lost over a step of length `step` by an electron
 * (qel 0) or positron (qel 1) of kinetic energy eke with range `range` */
float oracle_compute_eloss(const gegs_media_tables *mt, int qel, int medium, float range, float rhof, float step,
                           float eke) {
    float elke = om_logf(eke);
    gegs_float2 ek = mt->eke01[medium];
    int lelke = (int)(ek.x + ek.y * elke);
    return compute_eloss(mt, qel, medium, range, rhof, step, eke, elke, lelke);
}

/* range of an electron of kinetic energy eke, as electr_step forms it
 * (electron.cu:1414-1432): drange to the bin's lower edge plus range_ep */
float oracle_range(const gegs_media_tables *mt, int qel, int medium, float rhof, float eke) {
    float elke = om_logf(eke);
    gegs_float2 ek = mt->eke01[medium];
    int lelke = (int)(ek.x + ek.y * elke);
    const gegs_float2 *dedx = mt->dedx + qel * MXEKE * MXMED + medium * MXEKE;
    float ekei = mt->e_array[medium * MXEKE + lelke - 1];
    float elkei = ((float)lelke - ek.x) / ek.y;
    float range = drange(dedx, eke, elke, ekei, elkei, lelke);
    return (range + mt->range_ep[(medium * MXEKE + lelke - 1) * 2 + qel]) / rhof;
}

/* relax probe: the cascade of a vacancy in `shell` of element iz with energy
 * `energy`, in a region with the given record; returns the particles created
 * (at most max_out) and the energy deposited locally */
int oracle_relax(const gegs_media_tables *mt, const gegs_region *rd, int iz, int shell, float energy, uint64_t seed,
                 uint64_t lin, gegs_particle *out, int max_out, float *edep_out) {
    octx *c = (octx *)calloc(1, sizeof(octx));
    c->mt = mt;
    opart p;
    memset(&p, 0, sizeof(p));
    p.wt = 1.0f;
    p.region = VAC_AND_OTH;
    p.lin = lin;
    ostream s;
    st_begin(&s, &p, seed);
    float edep = 0.0f;
    relax(c, &s, energy, shell, iz, &edep, p, rd);
    int n = c->sp < max_out ? c->sp : max_out;
    for (int i = 0; i < n; ++i) {
        const opart *q = &c->stack[i];
        gegs_particle *o = out + i;
        memset(o, 0, sizeof(*o));
        o->e = q->e; o->u = q->u; o->v = q->v; o->w = q->w; o->wt = q->wt;
        o->charge = q->charge; o->status = q->status; o->lin = q->lin; o->blk = q->blk;
    }
    *edep_out = edep;
    free(c);
    return n;
}

/* EII shell choice probe: for n electrons of kinetic energy ekin in `medium`,
 * streams keyed by lineage lin0 + i: the chosen shell (1 K .. 4 L3) and
 * element, or 0 for a Moller interaction */
void oracle_eii_shell(const gegs_media_tables *mt, int medium, float ekin, uint64_t seed, uint64_t lin0, uint32_t n,
                      int32_t *shell, int32_t *z) {
    for (uint32_t i = 0; i < n; ++i) {
        opart p;
        memset(&p, 0, sizeof(p));
        p.lin = lin0 + i;
        ostream s;
        st_begin(&s, &p, seed);
        int ish = 0, iz = 0;
        float uj;
        if (!eii_choose(mt, &s, medium, ekin, &ish, &iz, &uj)) ish = iz = 0;
        shell[i] = ish;
        z[i] = iz;
    }
}

float oracle_logf(float x) { return om_logf(x); }
float oracle_expf(float x) { return om_expf(x); }
void oracle_sincos2pi(float t, float *s, float *c) { om_sincos2pi(t, s, c); }
void oracle_philox(uint32_t *ctr, uint32_t k0, uint32_t k1) { philox(ctr, k0, k1); }

/* geometry probes for the index-parity tests */
uint32_t oracle_howfar_phantom(uint32_t nx, uint32_t ny, uint32_t nz, const float *xb, const float *yb,
                               const float *zb, const gegs_field *field, const float *pos_dir,
                               uint32_t region, float *t) {
    octx c;
    memset(&c, 0, sizeof(c));
    c.xb = xb; c.yb = yb; c.zb = zb; c.nx = nx; c.ny = ny; c.nz = nz; c.f = *field;
    opart p;
    memset(&p, 0, sizeof(p));
    p.x = pos_dir[0]; p.y = pos_dir[1]; p.z = pos_dir[2];
    p.u = pos_dir[3]; p.v = pos_dir[4]; p.w = pos_dir[5];
    p.region = region;
    return howfar_phantom(&c, &p, t);
}

float oracle_hownear_phantom(uint32_t nx, uint32_t ny, uint32_t nz, const float *xb, const float *yb,
                             const float *zb, const gegs_field *field, const float *pos,
                             uint32_t region) {
    octx c;
    memset(&c, 0, sizeof(c));
    c.xb = xb; c.yb = yb; c.zb = zb; c.nx = nx; c.ny = ny; c.nz = nz; c.f = *field;
    c.fixes = g_fixes;
    opart p;
    memset(&p, 0, sizeof(p));
    p.x = pos[0]; p.y = pos[1]; p.z = pos[2];
    p.region = region;
    return hownear_phantom(&c, &p);
}

/* the head -> phantom frame of one point and direction (to_phantom):
 * out = phantom x, y, z, u, v, w */
void oracle_to_phantom(const gegs_field *field, const float *pos_dir, float *out) {
    octx c;
    memset(&c, 0, sizeof(c));
    c.f = *field;
    c.fixes = g_fixes;
    opart p;
    memset(&p, 0, sizeof(p));
    p.x = pos_dir[0]; p.y = pos_dir[1]; p.z = pos_dir[2];
    p.u = pos_dir[3]; p.v = pos_dir[4]; p.w = pos_dir[5];
    to_phantom(&c, &p, out);
}

/* block / wedge probe for the geometry tests: howfar (returns the region,
 * updates t and *module) and hownear of one particle in the block (is_wedge
 * 0) or the wedge (1) of bw, with the module order of bw alone */
uint32_t oracle_howfar_block_wedge(const gegs_block_wedge *bw, int is_wedge, const gegs_particle *in, int32_t *module,
                                   float *t, float *tperp) {
    octx c;
    memset(&c, 0, sizeof(c));
    set_module_order(&c, NULL, bw);
    opart p;
    memset(&p, 0, sizeof(p));
    p.x = in->x; p.y = in->y; p.z = in->z; p.u = in->u; p.v = in->v; p.w = in->w;
    p.region = in->region; p.module = is_wedge ? GEGS_M_WEDGE : GEGS_M_BLOCK;
    *tperp = hownear(&c, &p);
    uint32_t r = howfar(&c, &p, t);
    *module = p.module;
    return r;
}
