// gegs_capi.hip -- C ABI of the transport library (include/gegs.h).
//
// Replaces the host <-> device contract of the reference (launch_kernels /
// runSimulation, CUDA_EGS.cu:226-900; uploads in copy.cu, init.cu:909-1051,
// media.cu:1030-1048): state lives in a context object, every call returns a
// status code, nothing calls exit().
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/gegs.h"
#include "ctimport.hpp"
#include "event.hpp"
#include "plain.hpp"

using namespace gegs;

struct gegs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    gegs_media_tables *d_media = nullptr;
    uint16_t *d_rcls = nullptr;          // class index per region
    gegs_region *d_rtab = nullptr;       // distinct region records
    uint32_t nclass = 0;
    uint32_t max_med = 0;                // largest non-vacuum medium index of the phantom
    uint32_t nmed = 0;                   // media in the loaded tables
    float *d_xb = nullptr, *d_yb = nullptr, *d_zb = nullptr;
    uint32_t nx = 0, ny = 0, nz = 0;
    size_t nvox = 0;
    gegs_field field{};
    gegs_field *d_fields = nullptr;      // beam geometry per job of the current run
    uint32_t *d_job_plan = nullptr;      // dose buffer per job
    size_t fields_cap = 0, job_plan_cap = 0;
    uint64_t hist_len = 0;               // per-history slots of the last run
    gegs_mlc *d_mlc = nullptr;           // MLC geometry
    gegs_mlc h_mlc{};
    bool have_mlc = false;
    std::vector<float> leaf_pos;         // 120 per leaf set
    gegs_float2 *d_endnodes = nullptr;   // leaf ends per job
    size_t endnodes_cap = 0;
    gegs_jaws h_jaws{};                  // jaw geometry
    bool have_jaws = false;
    std::vector<float> jaw_pos;          // x1, x2, y1, y2 per set
    gegs_float2 *d_jaws = nullptr;       // jaw polygons per job
    gegs_block_wedge h_bw{};             // block and wedge geometry
    gegs_float2 *d_bw = nullptr;         // their 18 vertices on the device
    uint32_t fixes = 0;                  // GEGS_FIX_* mask
    uint32_t media_opts = 0;             // ENV_PRESTA2 / ENV_EXACT_BCA of the loaded tables
    size_t jaws_cap = 0;
    uint32_t *d_beamlet = nullptr;       // beamlet of each history slot
    float *d_bweight = nullptr;          // weight of each beamlet
    double *d_adj = nullptr, *d_grad = nullptr;
    size_t beamlet_cap = 0, bweight_cap = 0, adj_cap = 0, grad_cap = 0;
    uint64_t n_slots_bl = 0;             // history slots covered by d_beamlet
    uint32_t n_beamlets = 0;
    bool have_media = false, have_phantom = false, have_field = false;
    gegs_source *d_src = nullptr;
    uint64_t nsrc = 0;
    double *d_chunk = nullptr, *d_dose = nullptr, *d_dose2 = nullptr;
    size_t tally_n = 0;                  // capacity of d_dose / d_dose2
    size_t chunk_n = 0;                  // capacity of d_chunk (batch rows)
    float *d_stack = nullptr;
    size_t stack_bytes = 0;
    uint32_t *d_next = nullptr;
    unsigned long long *d_counters = nullptr;
    bool hist_tally = false;
    double *d_hist_edep = nullptr;
    uint32_t *d_hist_steps = nullptr;
    uint64_t hist_n = 0;
    gegs_counters counters{};
    float last_ms = 0.0f;
    EventBanks banks;
    EventConfig tuning;
    hipError_t scratch_err = hipSuccess; // address of this device's no-alloc scratch
};

static int fail(gegs_ctx *c, int code, const char *what, hipError_t e = hipSuccess) {
    if (c) {
        c->err = what;
        if (e != hipSuccess) {
            c->err += ": ";
            c->err += hipGetErrorString(e);
        }
    }
    return code;
}

#define HIPCHECK(c, call, what)                                        \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return fail((c), GEGS_ERR_DEVICE, (what), e_); \
    } while (0)

// Device memory of gegs_run_plain_noalloc: one copy of this array per GPU,
// held from the library's load on that GPU until the process ends.  In host
// code the name `scratch` is not a device address; the address of each GPU's
// copy is taken once per device with hipGetSymbolAddress (gegs_create) and
// only that address is used.  Calls on one GPU share its copy: concurrent
// no-alloc runs on one device race.
static __device__ __align__(256) unsigned char scratch[GEGS_NOALLOC_SCRATCH_BYTES];
static constexpr int MAX_DEVICES = 64;
static std::mutex scratch_mutex;
static unsigned char *scratch_addr[MAX_DEVICES];   // per device, nullptr until resolved

// address of this device's scratch copy (the device must be current)
static hipError_t resolve_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (scratch_addr[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    scratch_addr[device] = static_cast<unsigned char *>(ptr);
    return hipSuccess;
}

static unsigned char *scratch_of(int device) {
    std::lock_guard<std::mutex> lock(scratch_mutex);
    return (device >= 0 && device < MAX_DEVICES) ? scratch_addr[device] : nullptr;
}

template <class T>
static hipError_t realloc_dev(T *&p, size_t &have, size_t need_elems) {
    if (p && have >= need_elems) return hipSuccess;
    if (p) {
        hipError_t e = hipFree(p);
        if (e != hipSuccess) return e;
        p = nullptr;
    }
    have = 0;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), need_elems * sizeof(T));
    if (e == hipSuccess) have = need_elems;
    return e;
}

extern "C" {

size_t gegs_sizeof_media_tables(void) { return sizeof(gegs_media_tables); }

gegs_ctx *gegs_create(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    gegs_ctx *c = new gegs_ctx();
    c->device = device;
    // the no-alloc entry point reports a failure here when it is called
    c->scratch_err = resolve_scratch(device);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&c->d_next, 4096 * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&c->d_counters, 16 * sizeof(unsigned long long)) != hipSuccess) {
        delete c;
        return nullptr;
    }
    return c;
}

void gegs_destroy(gegs_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    hipFree(c->d_media);
    hipFree(c->d_rcls);
    hipFree(c->d_rtab);
    hipFree(c->d_xb);
    hipFree(c->d_yb);
    hipFree(c->d_zb);
    hipFree(c->d_src);
    hipFree(c->d_chunk);
    hipFree(c->d_dose);
    hipFree(c->d_dose2);
    hipFree(c->d_stack);
    hipFree(c->d_next);
    hipFree(c->d_counters);
    hipFree(c->d_hist_edep);
    hipFree(c->d_hist_steps);
    hipFree(c->d_fields);
    hipFree(c->d_job_plan);
    hipFree(c->d_mlc);
    hipFree(c->d_endnodes);
    hipFree(c->d_jaws);
    hipFree(c->d_bw);
    hipFree(c->d_beamlet);
    hipFree(c->d_bweight);
    hipFree(c->d_adj);
    hipFree(c->d_grad);
    c->banks.release();
    hipStreamDestroy(c->stream);
    delete c;
}

const char *gegs_last_error(gegs_ctx *c) { return c ? c->err.c_str() : "no context"; }

int gegs_load_media(gegs_ctx *c, const gegs_media_tables *t) {
    if (!c || !t) return fail(c, GEGS_ERR_ARG, "gegs_load_media: null argument");
    if (t->nmed < 1 || t->nmed > GEGS_MXMED) return fail(c, GEGS_ERR_MEDIA, "gegs_load_media: nmed out of range");
    // the reference drops in-flight annihilations when nbr_split <= 0 (quirk
    // Q5, annih.cu:71-75) and loses their energy: such tables are refused
    if (t->nbr_split < 1) return fail(c, GEGS_ERR_MEDIA, "gegs_load_media: nbr_split must be >= 1 (quirk Q5)");
    if (t->spin_effects) return fail(c, GEGS_ERR_MEDIA, "gegs_load_media: spin effects are not built (no spin data)");
    if (t->transport_algorithm != 0 && t->transport_algorithm != 1)
        return fail(c, GEGS_ERR_MEDIA, "gegs_load_media: transport_algorithm is 0 (PRESTA-II) or 1 (PRESTA-I)");
    hipSetDevice(c->device);
    if (!c->d_media) HIPCHECK(c, hipMalloc(&c->d_media, sizeof(gegs_media_tables)), "hipMalloc media");
    HIPCHECK(c, hipMemcpyAsync(c->d_media, t, sizeof(gegs_media_tables), hipMemcpyHostToDevice, c->stream), "copy media");
    HIPCHECK(c, hipStreamSynchronize(c->stream), "sync media");
    c->nmed = (uint32_t)t->nmed;
    c->media_opts = (t->transport_algorithm == 0 ? ENV_PRESTA2 : 0u) | (t->exact_bca ? ENV_EXACT_BCA : 0u);
    c->have_media = true;
    return GEGS_OK;
}

int gegs_validate_regions(const gegs_region *regions, uint64_t n) {
    if (!regions) return GEGS_ERR_ARG;
    for (uint64_t i = 0; i < n; ++i) {
        const uint16_t m = regions[i].med;
        if (m != GEGS_VACUUM && m >= GEGS_MXMED) return GEGS_ERR_PHANTOM;
        if (regions[i].flags & ~GEGS_F_SUPPORTED) return GEGS_ERR_ARG;
    }
    return GEGS_OK;
}

int gegs_load_phantom(gegs_ctx *c, uint32_t nx, uint32_t ny, uint32_t nz, const float *xb, const float *yb,
                      const float *zb, const gegs_region *regions) {
    if (!c || !xb || !yb || !zb || !regions) return fail(c, GEGS_ERR_ARG, "gegs_load_phantom: null argument");
    if (nx < 2 || ny < 2 || nz < 2) return fail(c, GEGS_ERR_PHANTOM, "gegs_load_phantom: need at least 2 voxels per axis");
    if (nx > GEGS_MAX_AXIS || ny > GEGS_MAX_AXIS || nz > GEGS_MAX_AXIS)
        return fail(c, GEGS_ERR_PHANTOM, "gegs_load_phantom: more than 4096 voxels along an axis");
    const size_t nvox = (size_t)nx * ny * nz;
    if (nvox + VAC_AND_OTH > 0xFFFFFFFFull) return fail(c, GEGS_ERR_PHANTOM, "gegs_load_phantom: too many voxels");
    // voxel media must exist in the loaded tables; region flags must name physics the kernels implement
    const int vrc = gegs_validate_regions(regions, nvox + VAC_AND_OTH);
    if (vrc == GEGS_ERR_PHANTOM) return fail(c, vrc, "gegs_load_phantom: region medium out of range");
    if (vrc) return fail(c, vrc, "gegs_load_phantom: region flags request physics this build does not implement");
    const float *b[3] = {xb, yb, zb};
    const uint32_t n[3] = {nx, ny, nz};
    float **d[3] = {&c->d_xb, &c->d_yb, &c->d_zb};
    hipSetDevice(c->device);
    for (int a = 0; a < 3; ++a) {
        for (uint32_t i = 0; i < n[a]; ++i)
            if (!(b[a][i + 1] > b[a][i])) return fail(c, GEGS_ERR_PHANTOM, "gegs_load_phantom: bounds must increase");
        // one extrapolated bound past the end keeps the neighbouring-voxel read
        // of the reference's hownear defined
        std::vector<float> pad(b[a], b[a] + n[a] + 1);
        pad.push_back(b[a][n[a]] + (b[a][n[a]] - b[a][n[a] - 1]));
        hipFree(*d[a]);
        *d[a] = nullptr;
        HIPCHECK(c, hipMalloc(d[a], pad.size() * sizeof(float)), "hipMalloc bounds");
        HIPCHECK(c, hipMemcpy(*d[a], pad.data(), pad.size() * sizeof(float), hipMemcpyHostToDevice), "copy bounds");
    }
    // deduplicate the region records: class index per region + record table
    const size_t nreg = nvox + VAC_AND_OTH;
    std::unordered_map<std::string, uint32_t> seen;
    std::vector<gegs_region> table;
    std::vector<uint16_t> cls(nreg);
    for (size_t i = 0; i < nreg; ++i) {
        const std::string key(reinterpret_cast<const char *>(&regions[i]), sizeof(gegs_region));
        auto it = seen.find(key);
        uint32_t k;
        if (it == seen.end()) {
            k = (uint32_t)table.size();
            if (k > 0xFFFFu) return fail(c, GEGS_ERR_PHANTOM, "gegs_load_phantom: more than 65536 distinct region records");
            seen.emplace(key, k);
            table.push_back(regions[i]);
        } else {
            k = it->second;
        }
        cls[i] = (uint16_t)k;
    }
    hipFree(c->d_rcls);
    hipFree(c->d_rtab);
    c->d_rcls = nullptr;
    c->d_rtab = nullptr;
    HIPCHECK(c, hipMalloc(&c->d_rcls, nreg * sizeof(uint16_t)), "hipMalloc region classes");
    HIPCHECK(c, hipMalloc(&c->d_rtab, table.size() * sizeof(gegs_region)), "hipMalloc region table");
    HIPCHECK(c, hipMemcpy(c->d_rcls, cls.data(), nreg * sizeof(uint16_t), hipMemcpyHostToDevice), "copy region classes");
    HIPCHECK(c, hipMemcpy(c->d_rtab, table.data(), table.size() * sizeof(gegs_region), hipMemcpyHostToDevice),
             "copy region table");
    c->nclass = (uint32_t)table.size();
    c->max_med = 0;
    for (const gegs_region &r : table)
        if (r.med != GEGS_VACUUM && r.med > c->max_med) c->max_med = r.med;
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    c->nvox = nvox;
    c->have_phantom = true;
    return GEGS_OK;
}

// modules a field may start in: the jaws, the MLC, the block, the wedge or the phantom
static bool valid_first(int32_t m) {
    return m == GEGS_M_PHANTOM || m == GEGS_M_MLC || m == GEGS_M_SECJAWY || m == GEGS_M_BLOCK || m == GEGS_M_WEDGE;
}

int gegs_set_field(gegs_ctx *c, const gegs_field *f) {
    if (!c || !f) return fail(c, GEGS_ERR_ARG, "gegs_set_field: null argument");
    if (!valid_first(f->module_first))
        return fail(c, GEGS_ERR_ARG, "gegs_set_field: fields start in the jaws, the MLC, the block, the wedge or the phantom");
    c->field = *f;
    c->have_field = true;
    return GEGS_OK;
}

int gegs_load_mlc(gegs_ctx *c, const gegs_mlc *m) {
    if (!c || !m) return fail(c, GEGS_ERR_ARG, "gegs_load_mlc: null argument");
    if (!(m->zmin > 0.0f) || !(m->zthick > 0.0f) || !(m->rmax > 0.0f) || !(m->x_end > m->x_start) ||
        !(m->wl_full > 0.0f) || !(m->wl_target > 0.0f) || !(m->wl_isocenter > 0.0f))
        return fail(c, GEGS_ERR_ARG, "gegs_load_mlc: inconsistent MLC geometry");
    hipSetDevice(c->device);
    if (!c->d_mlc) HIPCHECK(c, hipMalloc(&c->d_mlc, sizeof(gegs_mlc)), "hipMalloc MLC");
    HIPCHECK(c, hipMemcpy(c->d_mlc, m, sizeof(gegs_mlc), hipMemcpyHostToDevice), "copy MLC");
    c->h_mlc = *m;
    c->have_mlc = true;
    return GEGS_OK;
}

int gegs_load_jaws(gegs_ctx *c, const gegs_jaws *j) {
    if (!c || !j) return fail(c, GEGS_ERR_ARG, "gegs_load_jaws: null argument");
    if (!(j->rmax > 0.0f) || !(j->zmaxy > j->zminy) || !(j->zmaxx > j->zminx) || !(j->zminy > 0.0f) ||
        !(j->zminx > 0.0f))
        return fail(c, GEGS_ERR_ARG, "gegs_load_jaws: inconsistent jaw geometry");
    c->h_jaws = *j;
    c->have_jaws = true;
    return GEGS_OK;
}

int gegs_load_block_wedge(gegs_ctx *c, const gegs_block_wedge *bw) {
    if (!c || !bw) return fail(c, GEGS_ERR_ARG, "gegs_load_block_wedge: null argument");
    if (bw->has_wedge && (bw->wedge_nver < 3 || bw->wedge_nver > 6))
        return fail(c, GEGS_ERR_ARG, "gegs_load_block_wedge: the wedge cross-section needs 3 to 6 vertices");
    gegs_float2 v[18];
    for (int i = 0; i < 4; ++i) {
        v[i] = bw->block_xz[i];
        v[4 + i] = bw->block_yz[i];
        v[14 + i] = bw->wedge_bound[i];
    }
    for (int i = 0; i < 6; ++i) v[8 + i] = bw->wedge[i];
    hipSetDevice(c->device);
    if (!c->d_bw) HIPCHECK(c, hipMalloc(&c->d_bw, sizeof(v)), "hipMalloc block and wedge");
    HIPCHECK(c, hipMemcpy(c->d_bw, v, sizeof(v), hipMemcpyHostToDevice), "copy block and wedge");
    c->h_bw = *bw;
    return GEGS_OK;
}

int gegs_set_fixes(gegs_ctx *c, uint32_t mask) {
    if (!c) return fail(c, GEGS_ERR_ARG, "gegs_set_fixes: null context");
    if (mask & ~GEGS_FIX_ALL) return fail(c, GEGS_ERR_ARG, "gegs_set_fixes: unknown fix bits");
    c->fixes = mask;
    return GEGS_OK;
}

int gegs_set_jaw_positions(gegs_ctx *c, const float *xy, uint32_t n_sets) {
    if (!c || (!xy && n_sets)) return fail(c, GEGS_ERR_ARG, "gegs_set_jaw_positions: null argument");
    for (uint32_t k = 0; k < n_sets; ++k)
        if (!(xy[4 * k + 1] > xy[4 * k]) || !(xy[4 * k + 3] > xy[4 * k + 2]))
            return fail(c, GEGS_ERR_ARG, "gegs_set_jaw_positions: need x1 < x2 and y1 < y2");
    c->jaw_pos.assign(xy, xy + (size_t)n_sets * 4);
    return GEGS_OK;
}

int gegs_set_leaf_positions(gegs_ctx *c, const float *pos, uint32_t n_sets) {
    if (!c || (!pos && n_sets)) return fail(c, GEGS_ERR_ARG, "gegs_set_leaf_positions: null argument");
    c->leaf_pos.assign(pos, pos + (size_t)n_sets * 2 * GEGS_MLC_LEAVES);
    return GEGS_OK;
}

int gegs_set_source(gegs_ctx *c, const gegs_source *src, uint64_t n) {
    if (!c || (!src && n)) return fail(c, GEGS_ERR_ARG, "gegs_set_source: null argument");
    if (n > 0xFFFFFFF0ull) return fail(c, GEGS_ERR_ARG, "gegs_set_source: at most 2^32-16 records per call");
    hipSetDevice(c->device);
    hipFree(c->d_src);
    c->d_src = nullptr;
    c->nsrc = n;
    if (n == 0) return GEGS_OK;
    HIPCHECK(c, hipMalloc(&c->d_src, n * sizeof(gegs_source)), "hipMalloc source");
    HIPCHECK(c, hipMemcpy(c->d_src, src, n * sizeof(gegs_source), hipMemcpyHostToDevice), "copy source");
    return GEGS_OK;
}

int gegs_set_history_tally(gegs_ctx *c, int enable) {
    if (!c) return GEGS_ERR_ARG;
    c->hist_tally = enable != 0;
    return GEGS_OK;
}

// zeroed tallies for n_plans dose buffers and n_slots per-history slots
static int prepare_tallies(gegs_ctx *c, uint32_t n_plans, uint64_t n_slots, uint32_t batch_rows = 1) {
    const size_t n = c->nvox * n_plans;
    HIPCHECK(c, realloc_dev(c->d_chunk, c->chunk_n, n * batch_rows), "hipMalloc tally");
    size_t have = c->tally_n;
    HIPCHECK(c, realloc_dev(c->d_dose, have, n), "hipMalloc tally");
    have = c->tally_n;
    HIPCHECK(c, realloc_dev(c->d_dose2, have, n), "hipMalloc tally");
    if (c->tally_n < n) c->tally_n = n;
    const int blocks = 2048;
    hipLaunchKernelGGL(fill_zero, dim3(blocks), dim3(256), 0, c->stream, c->d_chunk, n * batch_rows);
    hipLaunchKernelGGL(fill_zero, dim3(blocks), dim3(256), 0, c->stream, c->d_dose, n);
    hipLaunchKernelGGL(fill_zero, dim3(blocks), dim3(256), 0, c->stream, c->d_dose2, n);
    hipLaunchKernelGGL(fill_zero_u64, dim3(1), dim3(64), 0, c->stream, c->d_counters, (size_t)16);
    hipLaunchKernelGGL(fill_zero_u32, dim3(16), dim3(256), 0, c->stream, c->d_next, (size_t)4096);
    c->hist_len = c->hist_tally ? n_slots : 0;
    if (c->hist_tally) {
        if (c->hist_n < n_slots) {
            hipFree(c->d_hist_edep);
            hipFree(c->d_hist_steps);
            c->d_hist_edep = nullptr;
            c->d_hist_steps = nullptr;
            HIPCHECK(c, hipMalloc(&c->d_hist_edep, n_slots * sizeof(double)), "hipMalloc history tally");
            HIPCHECK(c, hipMalloc(&c->d_hist_steps, n_slots * sizeof(uint32_t)), "hipMalloc history tally");
            c->hist_n = n_slots;
        }
        hipLaunchKernelGGL(fill_zero, dim3(1024), dim3(256), 0, c->stream, c->d_hist_edep, (size_t)n_slots);
        hipLaunchKernelGGL(fill_zero_u32, dim3(1024), dim3(256), 0, c->stream, c->d_hist_steps, (size_t)n_slots);
    }
    return GEGS_OK;
}

// leaf ends of one leaf set for a field (init_leaf_pos, init.cu:1594-1662):
// an open pair's bank A end moves to -pos[j+60], its bank B end to pos[j]
// (the banks are read inverted), scaled from the isocentre to the leaf
// centre plane; a closed pair keeps the closed profile
static void leaf_ends(const gegs_mlc &m, const float *pos, float sid, gegs_float2 *out) {
    const float zcenter = m.zmin + m.zthick / 2.0f;
    for (int j = 0; j < GEGS_MLC_LEAVES; ++j) {
        gegs_float2 *en = out + (size_t)j * 34;
        for (int i = 0; i < 34; ++i) en[i] = m.cross_yz[i];
        if (pos[j] + pos[j + GEGS_MLC_LEAVES] > 0.0f) {
            const float POS = pos[j], NEG = -pos[j + GEGS_MLC_LEAVES];
            for (int i = 2; i < 17; ++i) en[i].x = m.cross_yz[i].x + NEG * zcenter / sid;
            for (int i = 17; i < 32; ++i) en[i].x = m.cross_yz[i].x + POS * zcenter / sid;
        }
    }
}

// jaw polygons of one field (setup_secjaws, init.cu:1504-1592): the field
// rectangle at each slab's top (+-0.5 cm), the two slabs, and per pair the
// -side jaw, the opening and the +side jaw; layout as in jaws.hpp.  The
// +-0.5 offsets are double literals in the reference and are applied in
// double, the scalings in float.
static void jaw_polys(const gegs_jaws &g, const float *xy, float sid, gegs_float2 *out) {
    const float x1 = xy[0], x2 = xy[1], y1 = xy[2], y2 = xy[3];
    const float zmin[2] = {g.zminy, g.zminx}, zmax[2] = {g.zmaxy, g.zmaxx};
    for (int k = 0; k < 2; ++k) {
        gegs_float2 *b = out + 4 * k;
        const float a1 = x1 * zmin[k] / sid, a2 = x2 * zmin[k] / sid;
        const float c1 = y1 * zmin[k] / sid, c2 = y2 * zmin[k] / sid;
        b[0] = {(float)((double)a1 - 0.5), (float)((double)c1 - 0.5)};
        b[1] = {(float)((double)a2 + 0.5), (float)((double)c1 - 0.5)};
        b[2] = {(float)((double)a2 + 0.5), (float)((double)c2 + 0.5)};
        b[3] = {(float)((double)a1 - 0.5), (float)((double)c2 + 0.5)};
        gegs_float2 *m = out + 8 + 4 * k;
        m[0] = {-g.rmax, zmin[k]};
        m[1] = {g.rmax, zmin[k]};
        m[2] = {g.rmax, zmax[k]};
        m[3] = {-g.rmax, zmax[k]};
        // the Y pair's polygon uses the y openings, the X pair's the x openings
        const float lo = k == 0 ? y1 : x1, hi = k == 0 ? y2 : x2;
        gegs_float2 *e = out + 16 + 8 * k;
        e[0] = {-g.rmax, zmax[k]};
        e[1] = {-g.rmax, zmin[k]};
        e[2] = {lo * zmin[k] / sid, zmin[k]};
        e[3] = {lo * zmax[k] / sid, zmax[k]};
        e[4] = {hi * zmax[k] / sid, zmax[k]};
        e[5] = {hi * zmin[k] / sid, zmin[k]};
        e[6] = {g.rmax, zmin[k]};
        e[7] = {g.rmax, zmax[k]};
    }
}

// Head geometry of each job: jaw polygons (32 per job) of the fields that
// start in the jaws and leaf ends (60 x 34 per job) of the fields that pass
// the MLC.  With polys == ends == nullptr only validates and reports which
// of the two the jobs need.
static int job_geometry(gegs_ctx *c, const gegs_field *fields, size_t nf, const uint32_t *leaf_sets, size_t n_given,
                        bool &any_jaws, bool &any_mlc, gegs_float2 *polys, gegs_float2 *ends) {
    // a field starting in the jaws passes the MLC too when one is loaded
    auto uses_mlc = [&](const gegs_field &f) {
        return f.module_first == GEGS_M_MLC || (f.module_first == GEGS_M_SECJAWY && c->have_mlc);
    };
    any_jaws = any_mlc = false;
    for (size_t j = 0; j < nf; ++j) {
        const gegs_field &f = fields[j];
        any_mlc |= uses_mlc(f);
        any_jaws |= f.module_first == GEGS_M_SECJAWY;
        if (f.module_first == GEGS_M_BLOCK && !c->h_bw.has_block)
            return fail(c, GEGS_ERR_STATE, "run: a field starts in the block but no block is loaded");
        if (f.module_first == GEGS_M_WEDGE && !c->h_bw.has_wedge)
            return fail(c, GEGS_ERR_STATE, "run: a field starts in the wedge but no wedge is loaded");
    }
    if (any_jaws) {
        if (!c->have_jaws) return fail(c, GEGS_ERR_STATE, "run: a field starts in the jaws but no jaws are loaded");
        const size_t nsets = c->jaw_pos.size() / 4;
        for (size_t j = 0; j < nf; ++j) {
            if (fields[j].module_first != GEGS_M_SECJAWY) continue;
            const uint32_t set = j < n_given ? leaf_sets[j] : 0u;
            if (set >= nsets) return fail(c, GEGS_ERR_ARG, "run: a field's jaw set is not loaded");
            if (!(fields[j].source_to_isocenter > 0.0f)) return fail(c, GEGS_ERR_ARG, "run: source_to_isocenter must be > 0");
            if (polys) jaw_polys(c->h_jaws, c->jaw_pos.data() + set * 4, fields[j].source_to_isocenter, polys + j * 32);
        }
    }
    if (any_mlc) {
        if (!c->have_mlc) return fail(c, GEGS_ERR_STATE, "run: a field starts in the MLC but no MLC is loaded");
        const size_t nsets = c->leaf_pos.size() / (2 * GEGS_MLC_LEAVES);
        for (size_t j = 0; j < nf; ++j) {
            if (!uses_mlc(fields[j])) continue;
            const uint32_t set = j < n_given ? leaf_sets[j] : 0u;
            if (set >= nsets) return fail(c, GEGS_ERR_ARG, "run: a field's leaf set is not loaded");
            if (!(fields[j].source_to_isocenter > 0.0f)) return fail(c, GEGS_ERR_ARG, "run: source_to_isocenter must be > 0");
            if (ends)
                leaf_ends(c->h_mlc, c->leaf_pos.data() + set * 2 * GEGS_MLC_LEAVES, fields[j].source_to_isocenter,
                          ends + j * GEGS_MLC_LEAVES * 34);
        }
    }
    return GEGS_OK;
}

// beam geometry, dose buffer and leaf ends of each job, in device memory
static int upload_jobs(gegs_ctx *c, const std::vector<gegs_field> &fields, const std::vector<uint32_t> &plans,
                       const std::vector<uint32_t> &leaf_sets = {0u}) {
    bool any_jaws = false, any_mlc = false;
    int rc = job_geometry(c, fields.data(), fields.size(), leaf_sets.data(), leaf_sets.size(), any_jaws, any_mlc,
                          nullptr, nullptr);
    if (rc) return rc;
    std::vector<gegs_float2> polys(any_jaws ? fields.size() * 32 : 0);
    std::vector<gegs_float2> ends(any_mlc ? fields.size() * GEGS_MLC_LEAVES * 34 : 0);
    rc = job_geometry(c, fields.data(), fields.size(), leaf_sets.data(), leaf_sets.size(), any_jaws, any_mlc,
                      any_jaws ? polys.data() : nullptr, any_mlc ? ends.data() : nullptr);
    if (rc) return rc;
    HIPCHECK(c, realloc_dev(c->d_fields, c->fields_cap, fields.size()), "hipMalloc fields");
    HIPCHECK(c, realloc_dev(c->d_job_plan, c->job_plan_cap, plans.size()), "hipMalloc job plans");
    HIPCHECK(c, hipMemcpyAsync(c->d_fields, fields.data(), fields.size() * sizeof(gegs_field), hipMemcpyHostToDevice,
                               c->stream), "copy fields");
    HIPCHECK(c, hipMemcpyAsync(c->d_job_plan, plans.data(), plans.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                               c->stream), "copy job plans");
    if (any_jaws) {
        HIPCHECK(c, realloc_dev(c->d_jaws, c->jaws_cap, polys.size()), "hipMalloc jaw polygons");
        HIPCHECK(c, hipMemcpyAsync(c->d_jaws, polys.data(), polys.size() * sizeof(gegs_float2), hipMemcpyHostToDevice,
                                   c->stream), "copy jaw polygons");
    }
    if (any_mlc) {
        HIPCHECK(c, realloc_dev(c->d_endnodes, c->endnodes_cap, ends.size()), "hipMalloc leaf ends");
        HIPCHECK(c, hipMemcpyAsync(c->d_endnodes, ends.data(), ends.size() * sizeof(gegs_float2),
                                   hipMemcpyHostToDevice, c->stream), "copy leaf ends");
    }
    HIPCHECK(c, hipStreamSynchronize(c->stream), "copy jobs");
    return GEGS_OK;
}

static Env make_env(gegs_ctx *c, uint64_t seed, uint32_t njob) {
    Env env;
    env.mt = c->d_media;
    env.rcls = c->d_rcls;
    env.rtab = c->d_rtab;
    env.nclass = c->nclass;
    env.nmed = c->nmed;
    env.pge = c->d_media->ge;
    env.pgmfp = c->d_media->gmfp;
    env.pcohe = c->d_media->cohe;
    env.xb = c->d_xb;
    env.yb = c->d_yb;
    env.zb = c->d_zb;
    env.nx = c->nx;
    env.ny = c->ny;
    env.nz = c->nz;
    env.fields = c->d_fields;
    env.njob = njob;
    env.mlc = c->have_mlc ? c->d_mlc : nullptr;
    env.endnodes = c->d_endnodes;
    // module order jaws -> MLC -> block -> wedge -> phantom over the loaded modules (CUDA_EGS.cu:1279-1298)
    env.block_next = c->h_bw.has_wedge ? GEGS_M_WEDGE : GEGS_M_PHANTOM;
    env.mlc_next = c->h_bw.has_block ? GEGS_M_BLOCK : env.block_next;
    env.jaws = c->d_jaws;
    env.jawx_next = c->have_mlc ? GEGS_M_MLC : env.mlc_next;
    env.bw = c->d_bw;
    env.wedge_nver = c->h_bw.wedge_nver;
    env.fixes = c->fixes | c->media_opts;
    env.seed = seed;
    return env;
}

// dose and dose2 are host buffers, or device buffers with dev_out
static int finish_run(gegs_ctx *c, double *dose, double *dose2, hipEvent_t t0, hipEvent_t t1, uint32_t n_plans = 1,
                      bool dev_out = false) {
    HIPCHECK(c, hipEventSynchronize(t1), "run");
    hipEventElapsedTime(&c->last_ms, t0, t1);
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    unsigned long long h[16];
    HIPCHECK(c, hipMemcpy(h, c->d_counters, sizeof(h), hipMemcpyDeviceToHost), "read counters");
    c->counters.photon_steps = h[0];
    c->counters.electron_steps = h[1];
    c->counters.region_gathers = h[2];
    c->counters.deposits = h[3];
    for (int k = 0; k < 10; ++k) c->counters.interactions[k] = h[4 + k];
    c->counters.histories = h[14];
    c->counters.overflow = h[15];
    const size_t n = c->nvox * n_plans;
    if (dev_out) {
        if (dose)
            HIPCHECK(c, hipMemcpyAsync(dose, c->d_dose, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                     "copy dose");
        if (dose2)
            HIPCHECK(c, hipMemcpyAsync(dose2, c->d_dose2, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream),
                     "copy dose2");
        HIPCHECK(c, hipStreamSynchronize(c->stream), "copy dose");
    } else {
        if (dose) HIPCHECK(c, hipMemcpy(dose, c->d_dose, n * sizeof(double), hipMemcpyDeviceToHost), "read dose");
        if (dose2) HIPCHECK(c, hipMemcpy(dose2, c->d_dose2, n * sizeof(double), hipMemcpyDeviceToHost), "read dose2");
    }
    if (c->counters.overflow) return fail(c, GEGS_ERR_QUEUE_OVERFLOW, "particle stack overflow: particles were lost");
    return GEGS_OK;
}

static int check_ready(gegs_ctx *c, const gegs_run_options *o) {
    if (!c || !o) return fail(c, GEGS_ERR_ARG, "run: null argument");
    if (!c->have_media || !c->have_phantom || !c->have_field) return fail(c, GEGS_ERR_STATE, "run: load media, phantom and field first");
    if (c->nsrc == 0) return fail(c, GEGS_ERR_STATE, "run: no source particles");
    if (c->max_med >= c->nmed) return fail(c, GEGS_ERR_MEDIA, "run: the phantom uses a medium the loaded tables lack");
    if (o->n_batches < 1 || o->n_batches > 4096) return fail(c, GEGS_ERR_ARG, "run: n_batches must be in [1, 4096]");
    if (o->first >= c->nsrc || (o->count && o->first + o->count > c->nsrc))
        return fail(c, GEGS_ERR_ARG, "run: record range outside the loaded source");
    return GEGS_OK;
}

// the plain kernel instantiation for the record-table placement and the
// options of the run (the general kernel for any of ENV_GENERAL)
static void launch_plain(const PlainArgs &a, bool tab, uint32_t grid, uint32_t block, size_t lds, hipStream_t st) {
    const bool gen = (a.env.fixes & ENV_GENERAL) != 0;
    if (tab && gen) hipLaunchKernelGGL((plain_kernel<true, true>), dim3(grid), dim3(block), lds, st, a);
    else if (tab) hipLaunchKernelGGL((plain_kernel<true, false>), dim3(grid), dim3(block), lds, st, a);
    else if (gen) hipLaunchKernelGGL((plain_kernel<false, true>), dim3(grid), dim3(block), lds, st, a);
    else hipLaunchKernelGGL((plain_kernel<false, false>), dim3(grid), dim3(block), lds, st, a);
}

int gegs_run_plain(gegs_ctx *c, const gegs_run_options *o, double *dose, double *dose2) {
    int rc = check_ready(c, o);
    if (rc) return rc;
    hipSetDevice(c->device);
    rc = prepare_tallies(c, 1, c->nsrc);
    if (rc) return rc;
    rc = upload_jobs(c, {c->field}, {0u});
    if (rc) return rc;
    const uint32_t block = o->block_size ? o->block_size : 128;
    if (block != 128) return fail(c, GEGS_ERR_ARG, "run_plain: block_size must be 128");
    const uint32_t grid = o->grid_size ? o->grid_size : 2048;
    const int depth = o->stack_depth ? (int)o->stack_depth : 40;
    const size_t nthreads = (size_t)grid * block;
    size_t have = c->stack_bytes / sizeof(float);
    HIPCHECK(c, realloc_dev(c->d_stack, have, nthreads * (size_t)depth * PART_WORDS), "hipMalloc stack");
    c->stack_bytes = have * sizeof(float);
    PlainArgs a;
    a.env = make_env(c, o->seed, 1);
    a.src = c->d_src;
    a.hist_offset = o->history_offset;
    a.stack = c->d_stack;
    a.depth = depth;
    a.dose = c->d_chunk;
    a.hist_edep = c->hist_tally ? c->d_hist_edep : nullptr;
    a.hist_steps = c->hist_tally ? c->d_hist_steps : nullptr;
    a.counters = c->d_counters;
    const bool tab = c->nclass <= LDS_MAX_CLASSES;
    const size_t lds = env_lds_bytes(c->nx, c->ny, c->nz, c->nclass, c->nmed, 1, tab);
    hipEvent_t t0, t1;
    hipEventCreate(&t0);
    hipEventCreate(&t1);
    hipEventRecord(t0, c->stream);
    const uint32_t nb = o->n_batches;
    const uint64_t r0 = o->first, nr = o->count ? o->count : c->nsrc - o->first;
    for (uint32_t b = 0; b < nb; ++b) {
        a.first = (uint32_t)(r0 + (nr * b) / nb);
        a.last = (uint32_t)(r0 + (nr * (b + 1)) / nb);
        a.next = c->d_next + b;
        launch_plain(a, tab, grid, block, lds, c->stream);
        hipLaunchKernelGGL(fold_batch, dim3(2048), dim3(256), 0, c->stream, c->d_chunk, c->d_dose, c->d_dose2, c->nvox);
    }
    hipEventRecord(t1, c->stream);
    HIPCHECK(c, hipGetLastError(), "launch plain_kernel");
    return finish_run(c, dose, dose2, t0, t1);
}

// Layout of the no-alloc scratch for a run on nvox voxels: counters, batch
// work counters, the field, its plan index, jaw polygons, leaf ends, the
// batch dose row, then the particle stacks.
struct NoallocLayout {
    size_t counters = 0, next = 256, field, plan, polys, ends, chunk, stack;
    explicit NoallocLayout(size_t nvox) {
        auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
        field = up(next + 4096 * sizeof(uint32_t));
        plan = up(field + sizeof(gegs_field));
        polys = up(plan + sizeof(uint32_t));
        ends = up(polys + 32 * sizeof(gegs_float2));
        chunk = up(ends + (size_t)GEGS_MLC_LEAVES * 34 * sizeof(gegs_float2));
        stack = up(chunk + nvox * sizeof(double));
    }
};

int gegs_run_plain_noalloc(gegs_ctx *c, const gegs_run_options *o, double *dose, double *dose2) {
    int rc = check_ready(c, o);
    if (rc) return rc;
    if (!dose || !dose2) return fail(c, GEGS_ERR_ARG, "run_plain_noalloc: dose and dose2 are device buffers of nvox doubles");
    if (c->hist_tally) return fail(c, GEGS_ERR_STATE, "run_plain_noalloc: per-history tallies need gegs_run_plain");
    if (c->scratch_err != hipSuccess)
        return fail(c, GEGS_ERR_DEVICE, "run_plain_noalloc: address of the scratch array", c->scratch_err);
    HIPCHECK(c, hipSetDevice(c->device), "set device");
    unsigned char *base = scratch_of(c->device);
    if (!base) return fail(c, GEGS_ERR_DEVICE, "run_plain_noalloc: no scratch address for this device");
    const uint32_t block = o->block_size ? o->block_size : 128;
    if (block != 128) return fail(c, GEGS_ERR_ARG, "run_plain_noalloc: block_size must be 128");
    const int depth = o->stack_depth ? (int)o->stack_depth : 40;
    const NoallocLayout L(c->nvox);
    const size_t per_block = (size_t)block * depth * PART_WORDS * sizeof(float);
    const size_t cap = GEGS_NOALLOC_SCRATCH_BYTES;
    if (L.stack + per_block > cap)
        return fail(c, GEGS_ERR_ARG, "run_plain_noalloc: the phantom's dose row and one block of particle stacks "
                                     "exceed the scratch size (GEGS_NOALLOC_SCRATCH_BYTES)");
    const size_t fit = (cap - L.stack) / per_block;
    uint32_t grid = 2048;
    if (o->grid_size) {
        if (o->grid_size > fit)
            return fail(c, GEGS_ERR_ARG, "run_plain_noalloc: grid_size x stack_depth exceeds the scratch size");
        grid = o->grid_size;
    } else if (grid > fit) {
        grid = (uint32_t)fit;   // histories do not depend on the grid
    }
    // field, plan 0 and head geometry of leaf/jaw set 0, as gegs_run_plain
    bool any_jaws = false, any_mlc = false;
    gegs_float2 polys[32];
    gegs_float2 ends[GEGS_MLC_LEAVES * 34];
    rc = job_geometry(c, &c->field, 1, nullptr, 0, any_jaws, any_mlc, polys, ends);
    if (rc) return rc;
    const uint32_t plan0 = 0;
    HIPCHECK(c, hipMemcpyAsync(base + L.field, &c->field, sizeof(gegs_field), hipMemcpyHostToDevice, c->stream),
             "copy field");
    HIPCHECK(c, hipMemcpyAsync(base + L.plan, &plan0, sizeof(plan0), hipMemcpyHostToDevice, c->stream), "copy plan");
    if (any_jaws)
        HIPCHECK(c, hipMemcpyAsync(base + L.polys, polys, sizeof(polys), hipMemcpyHostToDevice, c->stream),
                 "copy jaw polygons");
    if (any_mlc)
        HIPCHECK(c, hipMemcpyAsync(base + L.ends, ends, sizeof(ends), hipMemcpyHostToDevice, c->stream),
                 "copy leaf ends");
    unsigned long long *counters = reinterpret_cast<unsigned long long *>(base + L.counters);
    uint32_t *next = reinterpret_cast<uint32_t *>(base + L.next);
    double *chunk = reinterpret_cast<double *>(base + L.chunk);
    // the scratch keeps the previous call's contents: clear what starts at zero
    hipLaunchKernelGGL(fill_zero_u64, dim3(1), dim3(64), 0, c->stream, counters, (size_t)16);
    hipLaunchKernelGGL(fill_zero_u32, dim3(16), dim3(256), 0, c->stream, next, (size_t)4096);
    hipLaunchKernelGGL(fill_zero, dim3(2048), dim3(256), 0, c->stream, chunk, c->nvox);
    hipLaunchKernelGGL(fill_zero, dim3(2048), dim3(256), 0, c->stream, dose, c->nvox);
    hipLaunchKernelGGL(fill_zero, dim3(2048), dim3(256), 0, c->stream, dose2, c->nvox);
    PlainArgs a;
    a.env = make_env(c, o->seed, 1);
    a.env.fields = reinterpret_cast<const gegs_field *>(base + L.field);
    a.env.jaws = reinterpret_cast<const gegs_float2 *>(base + L.polys);
    a.env.endnodes = reinterpret_cast<const gegs_float2 *>(base + L.ends);
    a.src = c->d_src;
    a.hist_offset = o->history_offset;
    a.stack = reinterpret_cast<float *>(base + L.stack);
    a.depth = depth;
    a.dose = chunk;
    a.hist_edep = nullptr;
    a.hist_steps = nullptr;
    a.counters = counters;
    const bool tab = c->nclass <= LDS_MAX_CLASSES;
    const size_t lds = env_lds_bytes(c->nx, c->ny, c->nz, c->nclass, c->nmed, 1, tab);
    const uint32_t nb = o->n_batches;
    const uint64_t r0 = o->first, nr = o->count ? o->count : c->nsrc - o->first;
    for (uint32_t b = 0; b < nb; ++b) {
        a.first = (uint32_t)(r0 + (nr * b) / nb);
        a.last = (uint32_t)(r0 + (nr * (b + 1)) / nb);
        a.next = next + b;
        launch_plain(a, tab, grid, block, lds, c->stream);
        hipLaunchKernelGGL(fold_batch, dim3(2048), dim3(256), 0, c->stream, chunk, dose, dose2, c->nvox);
    }
    HIPCHECK(c, hipGetLastError(), "launch plain_kernel");
    HIPCHECK(c, hipStreamSynchronize(c->stream), "run");
    unsigned long long h[16];
    HIPCHECK(c, hipMemcpy(h, counters, sizeof(h), hipMemcpyDeviceToHost), "read counters");
    c->counters.photon_steps = h[0];
    c->counters.electron_steps = h[1];
    c->counters.region_gathers = h[2];
    c->counters.deposits = h[3];
    for (int k = 0; k < 10; ++k) c->counters.interactions[k] = h[4 + k];
    c->counters.histories = h[14];
    c->counters.overflow = h[15];
    c->last_ms = 0.0f;
    c->hist_len = 0;
    if (c->counters.overflow) return fail(c, GEGS_ERR_QUEUE_OVERFLOW, "particle stack overflow: particles were lost");
    return GEGS_OK;
}

// One job of an event-queued run: a field, its dose buffer and its records.
struct JobSpec {
    uint32_t plan;
    uint64_t first, count, history_offset, slot0;
    uint32_t leaf_set;
};

// Event-queued transport of every job in one loop: each batch holds the
// b-th of n_batches equal slices of every job, so that every batch is an
// independent sample of every plan's dose.
// Batch rows held at once by a continuous run (bytes of per-batch dose rows).
static constexpr size_t CONTINUOUS_TALLY_BYTES = (size_t)16 << 30;

static int run_event_jobs(gegs_ctx *c, const gegs_run_options *o, const std::vector<gegs_field> &fields,
                          const std::vector<JobSpec> &jobs, uint32_t n_plans, uint64_t n_slots, double *dose,
                          double *dose2, const BeamletArgs &bl = BeamletArgs(), bool dev_out = false) {
    hipSetDevice(c->device);
    // Continuous batching: all statistical batches go through one event loop
    // (one drain at the end instead of one per batch), each deposit tallied
    // in the dose row of its history's batch; the rows are folded into dose
    // and dose2 afterwards.  The adjoint has no batch statistics and always
    // runs in one pass.  Falls back to one loop per batch when the rows do
    // not fit, or when a pass would exceed 2^32 - 16 histories.
    const uint32_t nb = o->n_batches;
    uint64_t total_all = 0;
    for (const JobSpec &js : jobs) total_all += js.count;
    const bool one_pass = total_all <= 0xFFFFFFF0ull &&
                          (bl.adj || (size_t)nb * n_plans * c->nvox * sizeof(double) <= CONTINUOUS_TALLY_BYTES);
    const uint32_t rows = (one_pass && !bl.adj) ? nb : 1u;
    int rc = prepare_tallies(c, n_plans, n_slots, rows);
    if (rc) return rc;
    std::vector<uint32_t> plans(jobs.size()), sets(jobs.size());
    for (size_t j = 0; j < jobs.size(); ++j) {
        plans[j] = jobs[j].plan;
        sets[j] = jobs[j].leaf_set;
    }
    rc = upload_jobs(c, fields, plans, sets);
    if (rc) return rc;
    const EventConfig cfg = c->tuning;
    Env env = make_env(c, o->seed, (uint32_t)jobs.size());
    hipError_t e = c->banks.reserve(total_all, nb, cfg);
    if (e != hipSuccess) return fail(c, GEGS_ERR_DEVICE, "event banks", e);
    hipLaunchKernelGGL(clear_overflow, dim3(1), dim3(1), 0, c->stream, c->banks.st);
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    (void)hipEventRecord(t0, c->stream);
    const size_t ntally = c->nvox * n_plans;
    const uint32_t passes = one_pass ? 1u : nb;
    for (uint32_t b = 0; b < passes; ++b) {
        BatchPlan &bp = *c->banks.h_plan;
        bp.njob = (uint32_t)jobs.size();
        uint64_t total = 0;
        for (size_t j = 0; j < jobs.size(); ++j) {
            const JobSpec &js = jobs[j];
            const uint64_t lo = js.count * b / passes, hi = js.count * (b + 1) / passes;
            bp.prefix[j] = (uint32_t)total;
            bp.rec0[j] = (uint32_t)(js.first + lo);
            bp.slot0[j] = (uint32_t)(js.slot0 + lo);
            bp.hist0[j] = js.history_offset + lo;
            total += hi - lo;
        }
        if (total > 0xFFFFFFF0ull) return fail(c, GEGS_ERR_ARG, "run: more than 2^32-16 histories in one batch");
        bp.prefix[jobs.size()] = (uint32_t)total;
        bp.total = (uint32_t)total;
        e = run_event_batch(c->banks, env, c->d_src, c->d_job_plan, c->nvox, c->d_chunk,
                            c->hist_tally ? c->d_hist_edep : nullptr, c->hist_tally ? c->d_hist_steps : nullptr,
                            c->d_counters, bl, cfg, c->stream, rows);
        if (e != hipSuccess) return fail(c, GEGS_ERR_DEVICE, "event batch", e);
        if (!bl.adj)
            hipLaunchKernelGGL(fold_batches, dim3(2048), dim3(256), 0, c->stream, c->d_chunk, c->d_dose, c->d_dose2,
                               c->nvox, (size_t)n_plans, rows);
    }
    hipLaunchKernelGGL(report_overflow, dim3(1), dim3(1), 0, c->stream, c->banks.st, c->d_counters);
    (void)hipEventRecord(t1, c->stream);
    HIPCHECK(c, hipGetLastError(), "launch event kernels");
    return finish_run(c, dose, dose2, t0, t1, n_plans, dev_out);
}

int gegs_run_event(gegs_ctx *c, const gegs_run_options *o, double *dose, double *dose2) {
    int rc = check_ready(c, o);
    if (rc) return rc;
    const uint64_t r0 = o->first, nr = o->count ? o->count : c->nsrc - o->first;
    // one job; per-history slots are record indices
    const JobSpec js{0u, r0, nr, o->history_offset + r0, r0, 0u};
    return run_event_jobs(c, o, {c->field}, {js}, 1, c->nsrc, dose, dose2);
}

static int parse_jobs(gegs_ctx *c, const gegs_job *jobs, uint32_t n_jobs, uint32_t n_plans, const gegs_run_options *o,
                      std::vector<gegs_field> &fields, std::vector<JobSpec> &specs, uint64_t &slots) {
    if (!c || !jobs || !o) return fail(c, GEGS_ERR_ARG, "run: null argument");
    if (n_jobs == 0 || n_jobs > MAX_JOBS) return fail(c, GEGS_ERR_ARG, "run: 1 to 512 jobs");
    if (n_plans == 0 || n_plans > n_jobs) return fail(c, GEGS_ERR_ARG, "run: 1 to n_jobs plans");
    if (!c->have_media || !c->have_phantom) return fail(c, GEGS_ERR_STATE, "run: load media and phantom first");
    if (c->nsrc == 0) return fail(c, GEGS_ERR_STATE, "run: no source particles");
    if (c->max_med >= c->nmed) return fail(c, GEGS_ERR_MEDIA, "run: the phantom uses a medium the loaded tables lack");
    if (o->n_batches < 1 || o->n_batches > 4096) return fail(c, GEGS_ERR_ARG, "run: n_batches must be in [1, 4096]");
    fields.resize(n_jobs);
    specs.resize(n_jobs);
    slots = 0;
    for (uint32_t j = 0; j < n_jobs; ++j) {
        const gegs_job &jb = jobs[j];
        if (jb.plan >= n_plans) return fail(c, GEGS_ERR_ARG, "run: job plan out of range");
        if (!valid_first(jb.field.module_first))
            return fail(c, GEGS_ERR_ARG, "run: fields start in the jaws, the MLC, the block, the wedge or the phantom");
        if (jb.count == 0 || jb.first >= c->nsrc || jb.first + jb.count > c->nsrc)
            return fail(c, GEGS_ERR_ARG, "run: job record range outside the loaded source");
        fields[j] = jb.field;
        specs[j] = JobSpec{jb.plan, jb.first, jb.count, jb.history_offset, slots, jb.leaf_set};
        slots += jb.count;
    }
    return GEGS_OK;
}

static int run_batched(gegs_ctx *c, const gegs_job *jobs, uint32_t n_jobs, uint32_t n_plans, const gegs_run_options *o,
                       double *dose, double *dose2, bool dev_out) {
    std::vector<gegs_field> fields;
    std::vector<JobSpec> specs;
    uint64_t slots = 0;
    int rc = parse_jobs(c, jobs, n_jobs, n_plans, o, fields, specs, slots);
    if (rc) return rc;
    BeamletArgs bl;
    if (c->d_beamlet) {
        if (c->n_slots_bl != slots) return fail(c, GEGS_ERR_ARG, "run: beamlet map does not cover the jobs' histories");
        bl.beamlet = c->d_beamlet;
        bl.bweight = c->d_bweight;
    }
    return run_event_jobs(c, o, fields, specs, n_plans, slots, dose, dose2, bl, dev_out);
}

int gegs_run_batched(gegs_ctx *c, const gegs_job *jobs, uint32_t n_jobs, uint32_t n_plans, const gegs_run_options *o,
                     double *dose, double *dose2) {
    return run_batched(c, jobs, n_jobs, n_plans, o, dose, dose2, false);
}

int gegs_run_batched_device(gegs_ctx *c, const gegs_job *jobs, uint32_t n_jobs, uint32_t n_plans,
                            const gegs_run_options *o, double *dose, double *dose2) {
    return run_batched(c, jobs, n_jobs, n_plans, o, dose, dose2, true);
}

int gegs_set_beamlets(gegs_ctx *c, const uint32_t *beamlet, uint64_t n_slots, const float *weight, uint32_t n_beamlets) {
    if (!c) return GEGS_ERR_ARG;
    hipSetDevice(c->device);
    if (!beamlet) {
        c->n_slots_bl = 0;
        c->n_beamlets = 0;
        hipFree(c->d_beamlet);
        hipFree(c->d_bweight);
        c->d_beamlet = nullptr;
        c->d_bweight = nullptr;
        c->beamlet_cap = c->bweight_cap = 0;
        return GEGS_OK;
    }
    if (!weight || n_slots == 0 || n_beamlets == 0) return fail(c, GEGS_ERR_ARG, "gegs_set_beamlets: bad argument");
    for (uint64_t k = 0; k < n_slots; ++k)
        if (beamlet[k] >= n_beamlets) return fail(c, GEGS_ERR_ARG, "gegs_set_beamlets: beamlet index out of range");
    HIPCHECK(c, realloc_dev(c->d_beamlet, c->beamlet_cap, n_slots), "hipMalloc beamlets");
    HIPCHECK(c, realloc_dev(c->d_bweight, c->bweight_cap, n_beamlets), "hipMalloc beamlet weights");
    HIPCHECK(c, hipMemcpy(c->d_beamlet, beamlet, n_slots * sizeof(uint32_t), hipMemcpyHostToDevice), "copy beamlets");
    HIPCHECK(c, hipMemcpy(c->d_bweight, weight, n_beamlets * sizeof(float), hipMemcpyHostToDevice), "copy weights");
    c->n_slots_bl = n_slots;
    c->n_beamlets = n_beamlets;
    return GEGS_OK;
}

// dLdD and grad are host buffers, or device buffers with dev_io
static int adjoint(gegs_ctx *c, const gegs_job *jobs, uint32_t n_jobs, uint32_t n_plans, const gegs_run_options *o,
                   const double *dLdD, double *grad, bool dev_io) {
    if (!c || !dLdD || !grad) return fail(c, GEGS_ERR_ARG, "gegs_adjoint: null argument");
    std::vector<gegs_field> fields;
    std::vector<JobSpec> specs;
    uint64_t slots = 0;
    int rc = parse_jobs(c, jobs, n_jobs, n_plans, o, fields, specs, slots);
    if (rc) return rc;
    if (!c->d_beamlet || c->n_slots_bl != slots)
        return fail(c, GEGS_ERR_STATE, "gegs_adjoint: set a beamlet map covering the jobs' histories first");
    hipSetDevice(c->device);
    const size_t n = c->nvox * n_plans;
    HIPCHECK(c, realloc_dev(c->d_adj, c->adj_cap, n), "hipMalloc dL/dD");
    HIPCHECK(c, realloc_dev(c->d_grad, c->grad_cap, c->n_beamlets), "hipMalloc gradient");
    // dL/dD goes to the context's buffer, whose address the captured sweeps keep
    HIPCHECK(c, hipMemcpyAsync(c->d_adj, dLdD, n * sizeof(double),
                               dev_io ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream), "copy dL/dD");
    hipLaunchKernelGGL(fill_zero, dim3(64), dim3(256), 0, c->stream, c->d_grad, (size_t)c->n_beamlets);
    BeamletArgs bl;
    bl.beamlet = c->d_beamlet;     // replay at unit beamlet weight: no bweight
    bl.adj = c->d_adj;
    bl.grad = c->d_grad;
    rc = run_event_jobs(c, o, fields, specs, n_plans, slots, nullptr, nullptr, bl);
    if (rc) return rc;
    HIPCHECK(c, hipMemcpyAsync(grad, c->d_grad, c->n_beamlets * sizeof(double),
                               dev_io ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream), "read gradient");
    HIPCHECK(c, hipStreamSynchronize(c->stream), "read gradient");
    return GEGS_OK;
}

int gegs_adjoint(gegs_ctx *c, const gegs_job *jobs, uint32_t n_jobs, uint32_t n_plans, const gegs_run_options *o,
                 const double *dLdD, double *grad) {
    return adjoint(c, jobs, n_jobs, n_plans, o, dLdD, grad, false);
}

int gegs_adjoint_device(gegs_ctx *c, const gegs_job *jobs, uint32_t n_jobs, uint32_t n_plans,
                        const gegs_run_options *o, const double *dLdD, double *grad) {
    return adjoint(c, jobs, n_jobs, n_plans, o, dLdD, grad, true);
}

int gegs_set_beamlet_weights_device(gegs_ctx *c, const float *weight, uint32_t n_beamlets) {
    if (!c || !weight) return fail(c, GEGS_ERR_ARG, "set_beamlet_weights_device: null argument");
    if (!c->d_beamlet || n_beamlets != c->n_beamlets)
        return fail(c, GEGS_ERR_STATE, "set_beamlet_weights_device: set a beamlet map with as many beamlets first");
    hipSetDevice(c->device);
    HIPCHECK(c, hipMemcpyAsync(c->d_bweight, weight, n_beamlets * sizeof(float), hipMemcpyDeviceToDevice, c->stream),
             "copy beamlet weights");
    HIPCHECK(c, hipStreamSynchronize(c->stream), "copy beamlet weights");
    return GEGS_OK;
}

int gegs_set_tuning(gegs_ctx *c, const gegs_tuning *t) {
    if (!c) return GEGS_ERR_ARG;
    EventConfig cfg;
    if (t) {
        if (t->block) cfg.block = t->block;
        if (t->flight_block) cfg.flight_block = t->flight_block;
        if (t->refill_min) cfg.refill_min = t->refill_min;
        if (t->grid) cfg.grid = t->grid;
        cfg.photon_grid = t->photon_grid;
        cfg.electron_grid = t->electron_grid;
        if (t->photon_chunk) cfg.photon_chunk = t->photon_chunk;
        if (t->electron_chunk) cfg.electron_chunk = t->electron_chunk;
        if (t->inject) cfg.inject = t->inject;
        if (t->threshold) cfg.threshold = t->threshold;
        if (t->sweeps_per_graph) cfg.sweeps_per_graph = t->sweeps_per_graph;
        if (t->capacity) cfg.capacity = (size_t)t->capacity;
        if (t->refill_slots) cfg.refill_slots = t->refill_slots;
        if (t->sweep_streams) cfg.sweep_streams = t->sweep_streams;
        if (t->tail_particles) cfg.tail = t->tail_particles == 0xFFFFFFFFu ? 0u : t->tail_particles;
    }
    if (cfg.tail > (1u << 18)) return fail(c, GEGS_ERR_ARG, "gegs_set_tuning: tail_particles above 262144");
    if (cfg.sweep_streams != 1 && cfg.sweep_streams != 2)
        return fail(c, GEGS_ERR_ARG, "gegs_set_tuning: sweep_streams must be 1 or 2");
    if (cfg.refill_slots < 64 || cfg.refill_slots > 4096 || cfg.refill_slots % 64)
        return fail(c, GEGS_ERR_ARG, "gegs_set_tuning: refill_slots must be a multiple of 64 in [64, 4096]");
    if (cfg.block != 256 && cfg.block != 128 && cfg.block != 64)
        return fail(c, GEGS_ERR_ARG, "gegs_set_tuning: block must be 64, 128 or 256");
    if (cfg.flight_block != 512 && cfg.flight_block != 256 && cfg.flight_block != 128 && cfg.flight_block != 64)
        return fail(c, GEGS_ERR_ARG, "gegs_set_tuning: flight_block must be 64, 128, 256 or 512");
    if (cfg.grid > 65536 || cfg.photon_grid > 65536 || cfg.electron_grid > 65536)
        return fail(c, GEGS_ERR_ARG, "gegs_set_tuning: grid above 65536 blocks");
    if (cfg.refill_min > 64) return fail(c, GEGS_ERR_ARG, "gegs_set_tuning: refill_min above 64");
    if (cfg.sweeps_per_graph > 64) return fail(c, GEGS_ERR_ARG, "gegs_set_tuning: sweeps_per_graph above 64");
    if (cfg.capacity > 0xFFFFFFF0ull) return fail(c, GEGS_ERR_ARG, "gegs_set_tuning: capacity above 2^32 - 16");
    if (cfg.capacity && cfg.capacity < (size_t)gegs::QSEG)
        return fail(c, GEGS_ERR_ARG, "gegs_set_tuning: capacity below one slot per append segment (32)");
    c->tuning = cfg;
    return GEGS_OK;
}

int gegs_get_tuning(gegs_ctx *c, gegs_tuning *t) {
    if (!c || !t) return GEGS_ERR_ARG;
    const EventConfig &cfg = c->tuning;
    t->block = cfg.block;
    t->grid = cfg.grid;
    t->photon_grid = cfg.photon_grid;
    t->electron_grid = cfg.electron_grid;
    t->photon_chunk = cfg.photon_chunk;
    t->electron_chunk = cfg.electron_chunk;
    t->inject = cfg.inject;
    t->threshold = cfg.threshold;
    t->sweeps_per_graph = cfg.sweeps_per_graph;
    t->flight_block = cfg.flight_block;
    t->refill_min = cfg.refill_min;
    t->refill_slots = cfg.refill_slots;
    t->sweep_streams = cfg.sweep_streams;
    t->tail_particles = cfg.tail ? cfg.tail : 0xFFFFFFFFu;
    t->capacity = cfg.capacity;
    return GEGS_OK;
}

int gegs_sample_events(gegs_ctx *c, const gegs_particle *in, uint32_t n, uint64_t seed, gegs_particle *out,
                       uint32_t max_out, uint32_t *n_out, double *edep) {
    if (!c || !in || !out || !n_out || !edep || n == 0 || max_out == 0)
        return fail(c, GEGS_ERR_ARG, "gegs_sample_events: bad argument");
    if (!c->have_media || !c->have_phantom || !c->have_field) return fail(c, GEGS_ERR_STATE, "gegs_sample_events: load first");
    if (c->max_med >= c->nmed) return fail(c, GEGS_ERR_MEDIA, "gegs_sample_events: the phantom uses a medium the loaded tables lack");
    for (uint32_t i = 0; i < n; ++i) {
        if (in[i].status < 0 || in[i].status >= N_EVENTS) return fail(c, GEGS_ERR_ARG, "gegs_sample_events: bad status");
        if ((size_t)in[i].region >= c->nvox + VAC_AND_OTH) return fail(c, GEGS_ERR_ARG, "gegs_sample_events: bad region");
    }
    hipSetDevice(c->device);
    gegs_particle *d_in = nullptr, *d_out = nullptr;
    uint32_t *d_n = nullptr;
    double *d_e = nullptr;
    HIPCHECK(c, hipMalloc(&d_in, n * sizeof(gegs_particle)), "hipMalloc sample");
    HIPCHECK(c, hipMalloc(&d_out, (size_t)n * max_out * sizeof(gegs_particle)), "hipMalloc sample");
    HIPCHECK(c, hipMalloc(&d_n, n * sizeof(uint32_t)), "hipMalloc sample");
    HIPCHECK(c, hipMalloc(&d_e, n * sizeof(double)), "hipMalloc sample");
    HIPCHECK(c, hipMemcpy(d_in, in, n * sizeof(gegs_particle), hipMemcpyHostToDevice), "copy sample");
    const int rc = upload_jobs(c, {c->field}, {0u});
    if (rc) return rc;
    Env env = make_env(c, seed, 1);
    hipLaunchKernelGGL(sample_kernel, dim3((n + 127) / 128), dim3(128), 0, c->stream, env, d_in, n, d_out, max_out, d_n, d_e);
    HIPCHECK(c, hipStreamSynchronize(c->stream), "sample kernel");
    HIPCHECK(c, hipMemcpy(out, d_out, (size_t)n * max_out * sizeof(gegs_particle), hipMemcpyDeviceToHost), "read sample");
    HIPCHECK(c, hipMemcpy(n_out, d_n, n * sizeof(uint32_t), hipMemcpyDeviceToHost), "read sample");
    HIPCHECK(c, hipMemcpy(edep, d_e, n * sizeof(double), hipMemcpyDeviceToHost), "read sample");
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    (void)hipFree(d_n);
    (void)hipFree(d_e);
    return GEGS_OK;
}

int gegs_history_tally(gegs_ctx *c, double *edep, uint32_t *steps) {
    if (!c) return GEGS_ERR_ARG;
    if (!c->hist_tally || !c->d_hist_edep) return fail(c, GEGS_ERR_STATE, "history tally not enabled");
    hipSetDevice(c->device);
    if (edep) HIPCHECK(c, hipMemcpy(edep, c->d_hist_edep, c->hist_len * sizeof(double), hipMemcpyDeviceToHost), "read edep");
    if (steps) HIPCHECK(c, hipMemcpy(steps, c->d_hist_steps, c->hist_len * sizeof(uint32_t), hipMemcpyDeviceToHost), "read steps");
    return GEGS_OK;
}

int gegs_set_kernel_timing(gegs_ctx *c, int enable) {
    if (!c) return GEGS_ERR_ARG;
    c->banks.timing = enable != 0;
    c->banks.photon_ms = 0.0;
    c->banks.photon_launches = 0;
    return GEGS_OK;
}

int gegs_kernel_timing(gegs_ctx *c, double *photon_ms, uint64_t *photon_launches) {
    if (!c || !photon_ms || !photon_launches) return GEGS_ERR_ARG;
    *photon_ms = c->banks.photon_ms;
    *photon_launches = c->banks.photon_launches;
    return GEGS_OK;
}

int gegs_atomic_probe(gegs_ctx *c, const uint32_t *idx, uint64_t n, uint32_t nvox, uint32_t reps, float *ms) {
    if (!c || !idx || !ms || n == 0 || nvox == 0 || reps == 0) return fail(c, GEGS_ERR_ARG, "gegs_atomic_probe: bad argument");
    for (uint64_t i = 0; i < n; ++i)
        if (idx[i] >= nvox) return fail(c, GEGS_ERR_ARG, "gegs_atomic_probe: index beyond nvox");
    hipSetDevice(c->device);
    double *acc = nullptr;
    uint32_t *d_idx = nullptr;
    HIPCHECK(c, hipMalloc(&acc, (size_t)nvox * sizeof(double)), "hipMalloc probe");
    hipError_t e = hipMalloc(&d_idx, n * sizeof(uint32_t));
    if (e != hipSuccess) {
        hipFree(acc);
        return fail(c, GEGS_ERR_DEVICE, "hipMalloc probe", e);
    }
    e = hipMemcpy(d_idx, idx, n * sizeof(uint32_t), hipMemcpyHostToDevice);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fill_zero, dim3(1024), dim3(256), 0, c->stream, acc, (size_t)nvox);
        hipLaunchKernelGGL(atomic_probe, dim3(2048), dim3(256), 0, c->stream, acc, d_idx, n, 1u);
        (void)hipEventRecord(t0, c->stream);
        hipLaunchKernelGGL(atomic_probe, dim3(2048), dim3(256), 0, c->stream, acc, d_idx, n, reps);
        (void)hipEventRecord(t1, c->stream);
        e = hipEventSynchronize(t1);
        if (e == hipSuccess) e = hipEventElapsedTime(ms, t0, t1);
    }
    if (t0) hipEventDestroy(t0);
    if (t1) hipEventDestroy(t1);
    hipFree(acc);
    hipFree(d_idx);
    if (e != hipSuccess) return fail(c, GEGS_ERR_DEVICE, "atomic probe", e);
    return GEGS_OK;
}

int gegs_ct_subsize(int32_t size[3], int16_t *hu, const float voxel[3], float offset[3], const float box_min[3],
                    const float box_max[3]) {
    if (!size || !hu || !voxel || !offset || !box_min || !box_max) return GEGS_ERR_ARG;
    for (int a = 0; a < 3; ++a)
        if (size[a] < 1 || size[a] > (int32_t)GEGS_MAX_CT_AXIS || !(voxel[a] > 0.0f)) return GEGS_ERR_ARG;
    int n[3] = {size[0], size[1], size[2]};
    gegs::ct::subsize(n, hu, voxel, offset, box_min, box_max);
    for (int a = 0; a < 3; ++a) size[a] = n[a];
    return GEGS_OK;
}

int gegs_ct_resample_grid(const int32_t size[3], const float voxel[3], float thickness[3], int32_t out_size[3]) {
    if (!size || !voxel || !thickness || !out_size) return GEGS_ERR_ARG;
    for (int a = 0; a < 3; ++a)
        if (size[a] < 1 || size[a] > (int32_t)GEGS_MAX_CT_AXIS || !(voxel[a] > 0.0f)) return GEGS_ERR_ARG;
    const int n[3] = {size[0], size[1], size[2]};
    int out[3];
    const int rc = gegs::ct::resample_grid(n, voxel, thickness, out, (int)GEGS_MAX_AXIS);
    if (rc) return rc;
    for (int a = 0; a < 3; ++a) out_size[a] = out[a];
    return GEGS_OK;
}

int gegs_ct_resample(const int32_t size[3], const float voxel[3], const int16_t *hu, const float offset[3],
                     const float thickness[3], const int32_t out_size[3], float *xb, float *yb, float *zb,
                     int16_t *out) {
    if (!size || !voxel || !hu || !offset || !thickness || !out_size || !xb || !yb || !zb || !out) return GEGS_ERR_ARG;
    for (int a = 0; a < 3; ++a) {
        if (size[a] < 1 || size[a] > (int32_t)GEGS_MAX_CT_AXIS || !(voxel[a] > 0.0f)) return GEGS_ERR_ARG;
        if (out_size[a] < 1 || out_size[a] > (int32_t)GEGS_MAX_AXIS || !(thickness[a] > 0.0f)) return GEGS_ERR_ARG;
    }
    const int n[3] = {size[0], size[1], size[2]};
    const int o[3] = {out_size[0], out_size[1], out_size[2]};
    gegs::ct::resample(n, voxel, hu, offset, thickness, o, xb, yb, zb, out);
    return GEGS_OK;
}

int gegs_get_counters(gegs_ctx *c, gegs_counters *out) {
    if (!c || !out) return GEGS_ERR_ARG;
    *out = c->counters;
    return GEGS_OK;
}

int gegs_last_kernel_ms(gegs_ctx *c, float *ms) {
    if (!c || !ms) return GEGS_ERR_ARG;
    *ms = c->last_ms;
    return GEGS_OK;
}

}  // extern "C"
