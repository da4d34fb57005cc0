// CPU build of the event-queued transport: the production kernels of
// gegs_amd/csrc/event.hpp, compiled as host C++ against the one-lane
// wavefront stand-in of include/hip/hip_runtime.h (or, with -DGEGS_SIMT,
// the 64-lane stand-in of include_simt/hip/hip_runtime.h, with the launch
// shape and refill policy of the GPU), run sweep after sweep until the
// banks and the source are empty.  Inputs and outputs are raw
// files in one directory (written and read by tests/test_event_cpu.py).
// Built with AddressSanitizer, with the LDS array sized exactly to what the
// flight kernels stage, so that any out-of-bounds access of the queue logic
// or of the per-particle code stops the run.
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../gegs_amd/csrc/event.hpp"

namespace gegs {
float lds_env[LDS_FLOATS];
}
using namespace gegs;

static std::vector<char> slurp(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) {
        fprintf(stderr, "missing %s\n", path.c_str());
        exit(2);
    }
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    std::vector<char> v((size_t)n);
    if (n && fread(v.data(), 1, (size_t)n, f) != (size_t)n) exit(3);
    fclose(f);
    return v;
}

template <class T> static T *copy_out(const std::vector<char> &v) {
    T *p = (T *)malloc(v.size());
    memcpy(p, v.data(), v.size());
    return p;
}

static void spill(const std::string &path, const void *p, size_t n) {
    FILE *f = fopen(path.c_str(), "wb");
    fwrite(p, 1, n, f);
    fclose(f);
}

struct Meta {
    uint32_t nx, ny, nz, nsrc, nbatch, pchunk, echunk, have_mlc, have_jaws, cap;
    uint64_t seed, hist_offset;
    // 64-lane build: flight block size, refill threshold and reservation,
    // blocks per flight launch, primaries per injection and the injection threshold
    uint32_t flight_block, refill_min, refill_slots, flight_grid, inject, threshold;
    // in-flight particles at or below which the tail launch finishes the run (0: never)
    uint32_t tail;
};

int main(int argc, char **argv) {
    if (argc < 2) return 1;
    const std::string d = argv[1];
    const std::vector<char> mv = slurp(d + "/meta.bin");
    Meta m;
    memcpy(&m, mv.data(), sizeof(m));
    const std::vector<char> tv = slurp(d + "/tables.bin");
    if (tv.size() != sizeof(gegs_media_tables)) return 4;
    const gegs_media_tables *mt = copy_out<gegs_media_tables>(tv);
    const size_t nvox = (size_t)m.nx * m.ny * m.nz, nreg = nvox + VAC_AND_OTH;
    const std::vector<char> rv = slurp(d + "/regions.bin");
    if (rv.size() != nreg * sizeof(gegs_region)) return 4;
    const gegs_region *regions = (const gegs_region *)rv.data();
    // bounds with one extrapolated entry, as gegs_load_phantom stores them
    std::vector<float> bounds[3];
    const char *names[3] = {"/xb.bin", "/yb.bin", "/zb.bin"};
    const uint32_t n[3] = {m.nx, m.ny, m.nz};
    for (int a = 0; a < 3; ++a) {
        const std::vector<char> v = slurp(d + names[a]);
        bounds[a].assign((const float *)v.data(), (const float *)v.data() + n[a] + 1);
        bounds[a].push_back(bounds[a][n[a]] + (bounds[a][n[a]] - bounds[a][n[a] - 1]));
    }
    // region classes, as gegs_load_phantom builds them
    std::map<std::string, uint32_t> seen;
    std::vector<gegs_region> table;
    std::vector<uint16_t> cls(nreg);
    for (size_t i = 0; i < nreg; ++i) {
        const std::string key((const char *)&regions[i], sizeof(gegs_region));
        auto it = seen.find(key);
        uint32_t k;
        if (it == seen.end()) {
            k = (uint32_t)table.size();
            seen.emplace(key, k);
            table.push_back(regions[i]);
        } else {
            k = it->second;
        }
        cls[i] = (uint16_t)k;
    }
    const std::vector<char> fv = slurp(d + "/field.bin");
    gegs_field field;
    memcpy(&field, fv.data(), sizeof(field));
    const std::vector<char> sv = slurp(d + "/src.bin");
    const gegs_source *src = (const gegs_source *)sv.data();
    const gegs_mlc *mlc = m.have_mlc ? copy_out<gegs_mlc>(slurp(d + "/mlc.bin")) : nullptr;
    const gegs_float2 *ends = m.have_mlc ? copy_out<gegs_float2>(slurp(d + "/endnodes.bin")) : nullptr;
    const gegs_float2 *jaws = m.have_jaws ? copy_out<gegs_float2>(slurp(d + "/jaws.bin")) : nullptr;

    Env env;
    env.mt = mt;
    env.rcls = cls.data();
    env.rtab = table.data();
    env.xb = bounds[0].data();
    env.yb = bounds[1].data();
    env.zb = bounds[2].data();
    env.pge = mt->ge;
    env.pgmfp = mt->gmfp;
    env.pcohe = mt->cohe;
    env.nx = m.nx;
    env.ny = m.ny;
    env.nz = m.nz;
    env.nclass = (uint32_t)table.size();
    env.nmed = (uint32_t)mt->nmed;
    env.fields = &field;
    env.njob = 1;
    env.mlc = mlc;
    env.endnodes = ends;
    env.mlc_next = GEGS_M_PHANTOM;
    env.jaws = jaws;
    env.jawx_next = mlc ? GEGS_M_MLC : GEGS_M_PHANTOM;
    env.seed = m.seed;
    const size_t lds = flight_lds(env);
    if (lds != sizeof(lds_env)) {
        fprintf(stderr, "LDS %zu bytes, array %zu\n", lds, sizeof(lds_env));
        return 7;
    }

    // the queue counters are 256-byte aligned blocks
    EvState *st = (EvState *)aligned_alloc(alignof(EvState), sizeof(EvState));
    memset(st, 0, sizeof(EvState));
    BatchPlan *bp = (BatchPlan *)calloc(1, sizeof(BatchPlan));
    uint32_t *job_plan = (uint32_t *)calloc(1, sizeof(uint32_t));
    // one dose row per statistical batch, all batches in one pass
    double *dose = (double *)calloc(nvox * m.nbatch, sizeof(double));
    double *hist_edep = (double *)calloc(m.nsrc, sizeof(double));
    uint32_t *hist_steps = (uint32_t *)calloc(m.nsrc, sizeof(uint32_t));
    unsigned long long *counters = (unsigned long long *)calloc(16, sizeof(unsigned long long));
    float *bank = (float *)calloc((size_t)m.cap * N_SLOTS * EVW, sizeof(float));
    EvArgs a;
    a.env = env;
    a.plan = bp;
    a.job_plan = job_plan;
    a.nvox = nvox;
    a.bank = bank;
    a.cap = m.cap;
    a.seg_cap = m.cap / QSEG;
    a.st = st;
    a.src = src;
    a.dose = dose;
    a.hist_edep = hist_edep;
    a.hist_steps = hist_steps;
    a.counters = counters;
    a.beamlet = nullptr;
    a.bweight = nullptr;
    a.adj = nullptr;
    a.grad = nullptr;
    a.tally_batches = m.nbatch;
    EventConfig cfg;
#ifdef GEGS_SIMT
    // 64-lane wavefronts with the GPU's launch shape and refill policy
    cfg.block = 256;
    cfg.grid = 4;
    cfg.flight_block = m.flight_block;
    cfg.photon_grid = m.flight_grid;
    cfg.electron_grid = m.flight_grid;
    cfg.refill_min = m.refill_min;
    cfg.refill_slots = m.refill_slots;
#else
    // one-lane wavefronts: a refill needs only one idle lane; the flight
    // kernels run one thread per block so that each stages all of LDS
    cfg.block = 64;
    cfg.grid = 2;
    cfg.flight_block = 1;
    cfg.photon_grid = 3;
    cfg.electron_grid = 3;
    cfg.refill_min = 1;
#endif
    cfg.photon_chunk = m.pchunk;
    cfg.electron_chunk = m.echunk;
    uint64_t sweeps = 0;
    bp->njob = 1;
    bp->total = m.nsrc;
    bp->prefix[0] = 0;
    bp->prefix[1] = m.nsrc;
    bp->rec0[0] = 0;
    bp->slot0[0] = 0;
    bp->hist0[0] = m.hist_offset;
    // a small injection size keeps several generations of particles in flight
    cfg.inject = m.inject ? m.inject : 256;
    cfg.threshold = m.threshold ? m.threshold : 512;
    cfg.tail = m.tail;
    float *tail_stack = m.tail ? (float *)calloc((size_t)tail_threads(cfg) * TAIL_DEPTH * PART_WORDS, sizeof(float))
                               : nullptr;
    cpu_launch(batch_begin, dim3(1), dim3(1), st, 0u, m.nsrc, cfg.threshold, cfg.inject);
    bool tailed = false;
    for (;;) {
        launch_sweep(a, cfg, nullptr);
        ++sweeps;
        if (st->overflow) return 5;
        if (!st->active) break;
        if (m.tail && st->src_next >= st->src_last && st->inj_n == 0 && st->inflight <= m.tail) {
            launch_tail(a, cfg, tail_stack, nullptr);
            tailed = true;
            if (st->overflow) return 5;
            if (st->active) return 7;
            break;
        }
        if (sweeps > 1000000) return 6;
    }
    spill(d + "/out_edep.bin", hist_edep, m.nsrc * sizeof(double));
    spill(d + "/out_steps.bin", hist_steps, m.nsrc * sizeof(uint32_t));
    spill(d + "/out_dose.bin", dose, nvox * m.nbatch * sizeof(double));
#ifdef GEGS_SIMT
    printf("sweeps %llu%s wavefront-ops %llu\n", (unsigned long long)sweeps, tailed ? " tail" : "",
           (unsigned long long)simt::S.ops);
#else
    printf("sweeps %llu%s\n", (unsigned long long)sweeps, tailed ? " tail" : "");
#endif
    return 0;
}
