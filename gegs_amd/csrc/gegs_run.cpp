// gegs_run -- a prepared run of the reference's input, transported from C++.
//
//   gegs_run <run file> [--device N]
//
// The run file (key = value lines, written by `python -m gegs_amd.run -i
// <file.ini> --prepare <run file>`) names the media-table file, the egsphant
// phantom, the head file and the phase-space files of the run input, with its
// field list and options.  This program does what the reference's main() and
// runSimulation() do from there (CUDA_EGS.cu:367-900, :1712-1735), through the
// library's C ABI only:
//   1. media tables, egsphant phantom and its region records, head geometry;
//   2. the phase-space headers, the batches of every field (nRepeat, batch
//      size, the last batch) and the records they load, decoded once per
//      distinct sequence and resident on the device;
//   3. the transport of every field: one event-queued loop over all fields
//      (batched, one dose buffer per field) or the plain kernel field after
//      field;
//   4. the plan's deposit (each field times its output factor), the dose file
//      <root>_DoseResults_NNNxNNNxNNN.bin and <root>_log.txt.
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/gegs.h"
#include "../../include/gegs_host.h"

namespace {

struct FieldSpec {
    uint64_t mu = 0;
    float angle = 0, cos_a = 1, sin_a = 0;
    bool has_of = false;
    float output_factor = 1;
};

struct RunFile {
    std::string dir, media, head, egsphant, phsp_base, root, method = "batched";
    char other[5][GEGS_NAME_LEN] = {};
    uint32_t flags = GEGS_F_REFERENCE_DEFAULT, phsp_files = 1, host_fixes = 0, device_fixes = 0, batches = 10;
    int32_t change_xyz = 0, module_first = GEGS_M_PHANTOM;
    float phsp_z = 0, sid = 100, iso[3] = {0, 0, 0}, dose_unit = 1.6e-8f;
    uint64_t histories_per_mu = 0, seed = 1;
    std::vector<FieldSpec> fields;
};

struct Log {
    FILE *f = nullptr;
    void operator()(const char *fmt, ...) {
        char buf[4096];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        fputs(buf, stdout);
        fputc('\n', stdout);
        fflush(stdout);
        if (f) {
            fputs(buf, f);
            fputc('\n', f);
            fflush(f);
        }
    }
};

std::string trim(const std::string &s) {
    const char *ws = " \t\r\n";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return std::string();
    return s.substr(a, s.find_last_not_of(ws) - a + 1);
}

[[noreturn]] void die(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    fputs("ERROR: ", stderr);
    vfprintf(stderr, fmt, ap);
    fputc('\n', stderr);
    va_end(ap);
    exit(code ? code : 1);
}

std::string resolve(const std::string &dir, const std::string &p) {
    if (p.empty() || p[0] == '/') return p;
    return dir + "/" + p;
}

RunFile read_run_file(const char *path) {
    RunFile r;
    FILE *f = fopen(path, "r");
    if (!f) die(2001, "Could not open the run file \"%s\".", path);
    const char *slash = strrchr(path, '/');
    r.dir = slash ? std::string(path, slash - path) : std::string(".");
    char line[8192];
    int n = 0;
    while (fgets(line, sizeof line, f)) {
        ++n;
        std::string s = trim(line);
        if (s.empty() || s[0] == '#') continue;
        const size_t eq = s.find('=');
        if (eq == std::string::npos) die(2001, "%s:%d: expected key = value", path, n);
        const std::string k = trim(s.substr(0, eq)), v = trim(s.substr(eq + 1));
        const char *c = v.c_str();
        if (k == "media file") r.media = resolve(r.dir, v);
        else if (k == "head file") r.head = resolve(r.dir, v);
        else if (k == "egsphant file") r.egsphant = resolve(r.dir, v);
        else if (k == "phsp base") r.phsp_base = resolve(r.dir, v);
        else if (k == "output root") r.root = resolve(r.dir, v);
        else if (k == "method") r.method = v;
        else if (k.rfind("region medium ", 0) == 0) {
            const int i = atoi(k.c_str() + 14);
            if (i < 1 || i > 5 || v.size() >= GEGS_NAME_LEN) die(2001, "%s:%d: bad region medium", path, n);
            strcpy(r.other[i - 1], v.c_str());
        } else if (k == "region flags") r.flags = (uint32_t)strtoul(c, nullptr, 0);
        else if (k == "phsp files") r.phsp_files = (uint32_t)strtoul(c, nullptr, 10);
        else if (k == "phsp z") r.phsp_z = strtof(c, nullptr);
        else if (k == "host fixes") r.host_fixes = (uint32_t)strtoul(c, nullptr, 0);
        else if (k == "device fixes") r.device_fixes = (uint32_t)strtoul(c, nullptr, 0);
        else if (k == "batches") r.batches = (uint32_t)strtoul(c, nullptr, 10);
        else if (k == "change xyz") r.change_xyz = atoi(c);
        else if (k == "module first") r.module_first = atoi(c);
        else if (k == "source to isocenter") r.sid = strtof(c, nullptr);
        else if (k == "isocenter") {
            if (sscanf(c, "%f %f %f", &r.iso[0], &r.iso[1], &r.iso[2]) != 3) die(2001, "%s:%d: isocenter", path, n);
        } else if (k == "dose unit") r.dose_unit = strtof(c, nullptr);
        else if (k == "histories per mu") r.histories_per_mu = strtoull(c, nullptr, 10);
        else if (k == "seed") r.seed = strtoull(c, nullptr, 10);
        else if (k == "field") {
            // MU, gantry angle (deg), cos and sin of the angle as the run input's
            // field geometry has them, output factor or "-"
            FieldSpec fs;
            unsigned long long mu;
            char of[64];
            if (sscanf(c, "%llu %f %f %f %63s", &mu, &fs.angle, &fs.cos_a, &fs.sin_a, of) != 5)
                die(2001, "%s:%d: field = <MU> <angle> <cos> <sin> <output factor | ->", path, n);
            fs.mu = mu;
            if (strcmp(of, "-") != 0) {
                fs.has_of = true;
                fs.output_factor = strtof(of, nullptr);
            }
            r.fields.push_back(fs);
        } else die(2001, "%s:%d: unknown key \"%s\"", path, n, k.c_str());
    }
    fclose(f);
    if (r.media.empty() || r.egsphant.empty() || r.phsp_base.empty() || r.root.empty() || r.fields.empty() ||
        !r.histories_per_mu)
        die(2001, "%s: media file, egsphant file, phsp base, output root, histories per mu and fields are required",
            path);
    for (int i = 0; i < 5; ++i)
        if (!r.other[i][0]) die(2001, "%s: region medium %d is missing", path, i + 1);
    bool any_of = false, all_of = true;
    for (const FieldSpec &fs : r.fields) {
        any_of |= fs.has_of;
        all_of &= fs.has_of;
    }
    if (any_of && !all_of) die(2001, "%s: output factors for some fields only", path);
    return r;
}

void check(gegs_ctx *ctx, int rc, const char *what) {
    if (rc) die(rc, "%s: %s", what, ctx ? gegs_last_error(ctx) : "");
}

void check_host(int rc, const char *what) {
    if (rc) die(rc, "%s: %s", what, gegs_host_last_error());
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) die(2001, "usage: gegs_run <run file> [--device N]");
    int device = 0;
    for (int i = 2; i < argc; ++i)
        if (!strcmp(argv[i], "--device") && i + 1 < argc) device = atoi(argv[++i]);
    const auto t0 = std::chrono::steady_clock::now();
    RunFile rf = read_run_file(argv[1]);
    Log log;
    log.f = fopen((rf.root + "_log.txt").c_str(), "w");
    if (!log.f) die(2001, "Could not open \"%s_log.txt\".", rf.root.c_str());
    log("GEGS (MI355X, C++ host) run file %s", argv[1]);

    // media tables
    std::vector<unsigned char> tbuf(sizeof(gegs_media_tables));
    gegs_media_tables *tables = reinterpret_cast<gegs_media_tables *>(tbuf.data());
    char names[GEGS_MXMED][GEGS_NAME_LEN];
    check_host(gegs_media_file_read(rf.media.c_str(), tables, names), "media tables");
    std::string media_list;
    for (int m = 0; m < tables->nmed; ++m) media_list += (m ? ", " : "") + std::string(names[m]);
    log("  media . . . . . . . . . . . %s", media_list.c_str());

    // phantom and region records
    gegs_egsphant ph;
    check_host(gegs_egsphant_read(rf.egsphant.c_str(), &ph), "egsphant");
    const size_t nvox = (size_t)ph.nx * ph.ny * ph.nz;
    std::vector<gegs_region> regions(nvox + 6);
    check_host(gegs_region_table(&ph, tables, names, rf.other, (uint16_t)rf.flags, regions.data()), "region table");
    std::vector<float> mass(nvox);
    check_host(gegs_voxel_mass(&ph, mass.data()), "voxel masses");
    log("  egsphant file . . . . . . . %s", rf.egsphant.c_str());
    log("  phantom voxels  . . . . . . %u x %u x %u", ph.nx, ph.ny, ph.nz);

    // head
    gegs_head head;
    memset(&head, 0, sizeof head);
    if (!rf.head.empty()) check_host(gegs_head_file_read(rf.head.c_str(), &head), "head file");
    const uint32_t nf = (uint32_t)rf.fields.size();
    if ((head.has_mlc || head.has_jaws) && head.n_sets != nf)
        die(2001, "the head file holds %u aperture sets for %u fields", head.n_sets, nf);

    // phase space and the batches of every field
    gegs_phsp *phsp = nullptr;
    check_host(gegs_phsp_open(rf.phsp_base.c_str(), rf.phsp_files, rf.host_fixes, &phsp), "phase space");
    const uint64_t total = gegs_phsp_total(phsp);
    for (uint32_t i = 0; i < rf.phsp_files || i == 0; ++i) {
        const char *hp = gegs_phsp_header_path(phsp, i);
        if (!hp) break;
        if (rf.phsp_files > 1) log("  Phase space file #%u . . . . %s", i, hp);
        else log("  Phase space file  . . . . . %s", hp);
    }
    log("  total particles . . . . . . %llu", (unsigned long long)total);
    std::vector<gegs_schedule> sched(nf);
    std::map<std::vector<uint64_t>, uint64_t> seq_first;   // load list -> first record in the source buffer
    std::vector<gegs_source> source;
    std::vector<uint64_t> first_of(nf);
    for (uint32_t i = 0; i < nf; ++i) {
        const uint64_t hist = rf.fields[i].mu * rf.histories_per_mu;
        check_host(gegs_field_schedule(hist, total, rf.host_fixes, &sched[i], nullptr, 0), "field schedule");
        std::vector<uint64_t> loads(2 * sched[i].n_loads);
        check_host(gegs_field_schedule(hist, total, rf.host_fixes, &sched[i], loads.data(), sched[i].n_loads),
                   "field schedule");
        auto it = seq_first.find(loads);
        if (it == seq_first.end()) {
            const uint64_t at = source.size();
            source.resize(at + sched[i].records);
            uint64_t k = at;
            for (uint64_t l = 0; l < sched[i].n_loads; ++l) {
                const int rc = gegs_phsp_read(phsp, loads[2 * l], loads[2 * l + 1], rf.phsp_z, source.data() + k);
                if (rc) die(rc == GEGS_PHSP_REFUSED ? 2019 : rc, "phase-space read of %llu records from %llu: %s",
                            (unsigned long long)loads[2 * l + 1], (unsigned long long)loads[2 * l],
                            rc == GEGS_PHSP_REFUSED ? "refused" : gegs_host_last_error());
                k += loads[2 * l + 1];
            }
            it = seq_first.emplace(loads, at).first;
        }
        first_of[i] = it->second;
    }
    gegs_phsp_close(phsp);

    // device
    gegs_ctx *ctx = gegs_create(device);
    if (!ctx) die(58001, "gegs_create(%d) failed", device);
    check(ctx, gegs_load_media(ctx, tables), "gegs_load_media");
    check(ctx, gegs_load_phantom(ctx, ph.nx, ph.ny, ph.nz, ph.xb, ph.yb, ph.zb, regions.data()), "gegs_load_phantom");
    if (head.has_mlc) {
        check(ctx, gegs_load_mlc(ctx, &head.mlc), "gegs_load_mlc");
        check(ctx, gegs_set_leaf_positions(ctx, head.leaf_positions, nf), "gegs_set_leaf_positions");
    }
    if (head.has_jaws) {
        check(ctx, gegs_load_jaws(ctx, &head.jaws), "gegs_load_jaws");
        check(ctx, gegs_set_jaw_positions(ctx, head.jaw_positions, nf), "gegs_set_jaw_positions");
    }
    if (head.has_block_wedge) check(ctx, gegs_load_block_wedge(ctx, &head.block_wedge), "gegs_load_block_wedge");
    check(ctx, gegs_set_fixes(ctx, rf.device_fixes), "gegs_set_fixes");
    check(ctx, gegs_set_source(ctx, source.data(), source.size()), "gegs_set_source");

    // jobs: every (field, repeat) pair with records, history indices in order
    std::vector<gegs_job> jobs;
    std::vector<uint32_t> job_field;
    uint64_t hist = 0, transported = 0;
    for (uint32_t i = 0; i < nf; ++i) {
        const FieldSpec &fs = rf.fields[i];
        const gegs_schedule &s = sched[i];
        gegs_field g;
        g.change_xyz = rf.change_xyz;
        g.module_first = rf.module_first;
        g.source_to_isocenter = rf.sid;
        g.cos_angle = fs.cos_a;
        g.sin_angle = fs.sin_a;
        g.iso_x = rf.iso[0];
        g.iso_y = rf.iso[1];
        g.iso_z = rf.iso[2];
        if (s.n_dropped)
            log("Field #%u: MU %llu, gantry %.1f, histories %llu, repeat times %u, batches %llu of %u, last %llu (%llu "
                "records not read, quirk Q16)",
                i, (unsigned long long)fs.mu, fs.angle, (unsigned long long)s.histories, s.n_repeat + 1,
                (unsigned long long)(s.n_batch + 1), s.batch_size, (unsigned long long)s.last_batch,
                (unsigned long long)s.dropped_records);
        else
            log("Field #%u: MU %llu, gantry %.1f, histories %llu, repeat times %u, batches %llu of %u, last %llu", i,
                (unsigned long long)fs.mu, fs.angle, (unsigned long long)s.histories, s.n_repeat + 1,
                (unsigned long long)(s.n_batch + 1), s.batch_size, (unsigned long long)s.last_batch);
        if (fs.has_of) log("  Output factor . . . . . . . %f", fs.output_factor);
        for (uint32_t q = 0; q <= s.n_repeat; ++q) {
            const uint64_t cnt = s.repeat_counts[q];
            if (!cnt) continue;
            gegs_job j;
            j.field = g;
            j.plan = i;
            j.leaf_set = i;
            j.first = first_of[i];
            j.count = cnt;
            j.history_offset = hist;
            jobs.push_back(j);
            job_field.push_back(i);
            hist += cnt;
            transported += cnt;
        }
    }

    std::vector<double> field_edep((size_t)nf * nvox, 0.0);
    gegs_run_options opt;
    memset(&opt, 0, sizeof opt);
    opt.seed = rf.seed;
    opt.n_batches = rf.batches;
    if (rf.method == "batched") {
        const size_t kMaxJobs = 512;
        for (size_t c0 = 0; c0 < jobs.size(); c0 += kMaxJobs) {
            const size_t c1 = std::min(jobs.size(), c0 + kMaxJobs);
            std::map<uint32_t, uint32_t> remap;   // field -> plan of this call, in field order
            for (size_t k = c0; k < c1; ++k) remap.emplace(job_field[k], 0);
            uint32_t np = 0;
            for (auto &kv : remap) kv.second = np++;
            std::vector<gegs_job> cj(jobs.begin() + c0, jobs.begin() + c1);
            for (size_t k = 0; k < cj.size(); ++k) cj[k].plan = remap[job_field[c0 + k]];
            std::vector<double> dose((size_t)np * nvox), dose2((size_t)np * nvox);
            check(ctx, gegs_run_batched(ctx, cj.data(), (uint32_t)cj.size(), np, &opt, dose.data(), dose2.data()),
                  "gegs_run_batched");
            gegs_counters cnt;
            check(ctx, gegs_get_counters(ctx, &cnt), "gegs_get_counters");
            if (cnt.overflow) die(90001, "particle banks overflowed");
            for (auto &kv : remap)
                for (size_t v = 0; v < nvox; ++v) field_edep[(size_t)kv.first * nvox + v] += dose[(size_t)kv.second * nvox + v];
        }
    } else if (rf.method == "plain") {
        std::vector<double> dose(nvox), dose2(nvox);
        for (size_t k = 0; k < jobs.size(); ++k) {
            const uint32_t i = job_field[k];
            if (head.has_mlc) check(ctx, gegs_set_leaf_positions(ctx, head.leaf_positions + (size_t)i * 2 * GEGS_MLC_LEAVES, 1), "leaves");
            if (head.has_jaws) check(ctx, gegs_set_jaw_positions(ctx, head.jaw_positions + (size_t)i * 4, 1), "jaws");
            check(ctx, gegs_set_field(ctx, &jobs[k].field), "gegs_set_field");
            // the plain kernel keys history k of its range by history_offset +
            // record index: the job's histories need the offset less the first
            // record (modulo 2^64, as the library adds it)
            opt.history_offset = jobs[k].history_offset - jobs[k].first;
            opt.first = jobs[k].first;
            opt.count = jobs[k].count;
            check(ctx, gegs_run_plain(ctx, &opt, dose.data(), dose2.data()), "gegs_run_plain");
            for (size_t v = 0; v < nvox; ++v) field_edep[(size_t)i * nvox + v] += dose[v];
        }
    } else {
        die(2001, "method must be batched or plain");
    }
    gegs_destroy(ctx);

    // the plan's deposit (CUDA_EGS.cu:1712-1735) and write_dose
    std::vector<double> edep(nvox, 0.0);
    for (uint32_t i = 0; i < nf; ++i) {
        const double of = (double)rf.fields[i].output_factor;
        const double *d = field_edep.data() + (size_t)i * nvox;
        if (rf.fields[i].has_of)
            for (size_t v = 0; v < nvox; ++v) edep[v] += d[v] * of;
        else
            for (size_t v = 0; v < nvox; ++v) edep[v] += d[v];
    }
    char fname[4200];
    check_host(gegs_write_dose(rf.root.c_str(), ph.nx, ph.ny, ph.nz, edep.data(), mass.data(), rf.dose_unit, nullptr,
                               fname, sizeof fname),
               "dose file");
    log("%s", fname);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    log("Total CPU/GPU . . . . . . . . %.2f ms (%.2f min)", secs * 1e3, secs / 60.0);
    printf("%llu histories in %.2f s -> %s\n", (unsigned long long)transported, secs, fname);
    gegs_head_free(&head);
    gegs_egsphant_free(&ph);
    fclose(log.f);
    return 0;
}
