// gegs_run -- a prepared run of the reference's input, transported from C++.
//
//   gegs_run -i <file.ini> --media <tables.gegsmed> [--base-dir D] [--method batched|plain]
//            [--seed S] [--batches B] [--fix Q1,Q4,Q7,Q8,Q9,Q10,Q15,Q16] [--dry-run <run file>] [--device N]
//   gegs_run <run file> [--device N]
//
// With -i it reads the reference's run input itself (CUDA_EGS.cu:936-1480):
// the [General], [Data], [Phase space], [SecJaws], [MLC], [Block], [Wedge],
// [CTcreate], [Phantom], [Calibration], [Output factor], [Interpolation] and
// [Dose output] keys with the reference's defaults and error codes, the
// .egsinp components (SECJAWS, DYNVMLC, PYRAMIDS), the .mlcinfo field list
// with its leaf files and the output-factor table; with "use ctcreate = Yes"
// it builds the phantom from the DICOM series and the CT ramps file and writes
// it to <output root>.egsphant (ctcreate.c:1342-1600); the physics tables come
// from a media-table file (the HATCH builders' output, like the reference's
// PEGS data: `python -m gegs_amd.hatch`).  --dry-run writes the run file and
// head file (and the CT phantom) instead of running.
//
// The run file (key = value lines) names the media-table file, the egsphant
// phantom, the head file and the phase-space files, with the field list and
// options.  From there this program does what the reference's main() and
// runSimulation() do (CUDA_EGS.cu:367-900, :1712-1735), through the library's
// C ABI only:
//   1. media tables, egsphant phantom and its region records, head geometry;
//   2. the phase-space headers, the batches of every field (nRepeat, batch
//      size, the last batch) and the records they load, decoded once per
//      distinct sequence and resident on the device;
//   3. the transport of every field: one event-queued loop over all fields
//      (batched, one dose buffer per field) or the plain kernel field after
//      field;
//   4. the plan's deposit (each field times its output factor), the dose file
//      <root>_DoseResults_NNNxNNNxNNN.bin and <root>_log.txt.
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

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
    bool head_in_memory = false;     // the -i form builds the head here instead of reading a head file
    gegs_head head_mem{};
    std::vector<float> leaf_mem, jaw_mem;
    char other[5][GEGS_NAME_LEN] = {};
    uint32_t flags = GEGS_F_REFERENCE_DEFAULT, phsp_files = 1, host_fixes = 0, device_fixes = 0, batches = 10;
    int32_t change_xyz = 0, module_first = GEGS_M_PHANTOM;
    float phsp_z = 0, sid = 100, iso[3] = {0, 0, 0}, dose_unit = 1.6e-8f;
    uint64_t histories_per_mu = 0, seed = 1;
    std::vector<FieldSpec> fields;
    // outputs of write_dose besides the dose file
    bool interp = false;             // interpolated dose: start (with the half voxel), sizes, points per axis
    float interp_start[3] = {0, 0, 0}, interp_sizes[3] = {0, 0, 0};
    uint32_t interp_dimen[3] = {0, 0, 0};
    bool absdose = false;            // absolute dose at a point, calibrated or not, against a measured dose
    float abs_pos[3] = {0, 0, 0}, abs_measured = 0;
    int abs_calibrated = 0;
    bool reldose = false;            // relative dose: points, the measured profile and the comparison
    float rel_start[3] = {0, 0, 0}, rel_sizes[3] = {0, 0, 0}, rel_iso[3] = {0, 0, 0}, rel_glen = 0, rel_gdose = 0;
    float rel_measured = 0;          // the absolute dose's measured value (a single point compares with it)
    uint32_t rel_dimen[3] = {0, 0, 0};
    int rel_norm = 0, rel_array = 0;
    std::string rel_file;            // measured profile
};

// logstr (output.c:41-49): every line to stdout and the log file; lines
// logged before the log file is open are written to it when it opens
struct Log {
    FILE *f = nullptr;
    std::vector<std::string> pending;
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
        } else {
            pending.push_back(buf);
        }
    }
    bool open(const std::string &path) {
        f = fopen(path.c_str(), "w");
        if (!f) return false;
        for (const std::string &l : pending) fprintf(f, "%s\n", l.c_str());
        pending.clear();
        fflush(f);
        return true;
    }
};

std::string trim(const std::string &s) {
    const char *ws = " \t\r\n";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return std::string();
    return s.substr(a, s.find_last_not_of(ws) - a + 1);
}

// the reference's error(): the code and the message, then the exit (status 1;
// the code itself does not fit an exit status)
[[noreturn]] void die(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "ERROR (%d): ", code);
    vfprintf(stderr, fmt, ap);
    fputc('\n', stderr);
    va_end(ap);
    exit(1);
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
        else if (k == "interpolation") {
            float *a = r.interp_start, *b = r.interp_sizes;
            uint32_t *d = r.interp_dimen;
            if (sscanf(c, "%f %f %f %f %f %f %u %u %u", &a[0], &a[1], &a[2], &b[0], &b[1], &b[2], &d[0], &d[1], &d[2]) != 9)
                die(2001, "%s:%d: interpolation = <start x y z> <sizes x y z> <points x y z>", path, n);
            r.interp = true;
        } else if (k == "absolute dose") {
            if (sscanf(c, "%f %f %f %d %f", &r.abs_pos[0], &r.abs_pos[1], &r.abs_pos[2], &r.abs_calibrated,
                       &r.abs_measured) != 5)
                die(2001, "%s:%d: absolute dose = <x y z> <calibrated 0|1> <measured>", path, n);
            r.absdose = true;
        } else if (k == "relative dose") {
            float *a = r.rel_start, *b = r.rel_sizes, *q = r.rel_iso;
            uint32_t *d = r.rel_dimen;
            char file[4096] = {0};
            if (sscanf(c, "%f %f %f %f %f %f %u %u %u %f %f %f %d %d %f %f %d %f %4095s", &a[0], &a[1], &a[2], &b[0], &b[1],
                       &b[2], &d[0], &d[1], &d[2], &q[0], &q[1], &q[2], &r.rel_norm, &r.rel_array, &r.rel_glen,
                       &r.rel_gdose, &r.abs_calibrated, &r.rel_measured, file) != 19)
                die(2001, "%s:%d: relative dose = <start> <sizes> <points> <isocentre> <norm> <array> <gamma cm> "
                          "<gamma fraction> <calibrated> <measured dose> <measured file>", path, n);
            r.rel_file = resolve(r.dir, file);
            r.reldose = true;
        }
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

// ------------------------------------------------------------------ paths (posix rules)
bool is_abs(const std::string &p) { return !p.empty() && p[0] == '/'; }

std::string join(const std::string &a, const std::string &b) {
    if (is_abs(b) || a.empty()) return b;
    return a.back() == '/' ? a + b : a + "/" + b;
}

// normalised path: "." and empty parts dropped, ".." folded where it can be
std::string normpath(const std::string &p) {
    if (p.empty()) return ".";
    const bool abs = is_abs(p);
    std::vector<std::string> parts;
    size_t i = 0;
    while (i <= p.size()) {
        size_t j = p.find('/', i);
        if (j == std::string::npos) j = p.size();
        const std::string c = p.substr(i, j - i);
        if (c.empty() || c == ".") {
        } else if (c == "..") {
            if (!parts.empty() && parts.back() != "..") parts.pop_back();
            else if (!abs) parts.push_back(c);
        } else {
            parts.push_back(c);
        }
        i = j + 1;
    }
    std::string out = abs ? "/" : "";
    for (size_t k = 0; k < parts.size(); ++k) out += (k ? "/" : "") + parts[k];
    return out.empty() ? "." : out;
}

std::string cwd() {
    char buf[4096];
    return getcwd(buf, sizeof buf) ? std::string(buf) : std::string(".");
}

std::string abspath(const std::string &p) { return normpath(is_abs(p) ? p : join(cwd(), p)); }

std::string slashes(std::string s) {
    for (char &c : s)
        if (c == '\\') c = '/';
    return s;
}

// a file name of the run input on this host: backslashes as separators,
// relative names against base (the reference resolves them with
// GetFullPathName against its working directory)
std::string host_path(const std::string &base, const std::string &name) {
    const std::string n = slashes(name);
    return is_abs(n) ? n : normpath(join(base, n));
}

bool exists(const std::string &p) {
    FILE *f = fopen(p.c_str(), "rb");
    if (!f) return false;
    fclose(f);
    return true;
}

// ------------------------------------------------------------------ the run input (.ini)
const char *kMissing = "<<missing>>";

// GetPrivateProfileString's reading: section and key names compare
// case-insensitively, blanks around them and around the value are dropped,
// one pair of enclosing double quotes is removed; a repeated key keeps its
// first value
struct Ini {
    std::map<std::string, std::map<std::string, std::string>> sec;
    std::string path;
    static std::string lower(std::string s) {
        for (char &c : s) c = (char)tolower((unsigned char)c);
        return s;
    }
    std::string get(const char *section, const char *key) const {
        auto a = sec.find(lower(section));
        if (a == sec.end()) return kMissing;
        auto b = a->second.find(lower(key));
        return b == a->second.end() ? std::string(kMissing) : b->second;
    }
    std::string req(const char *section, const char *key, int code) const {
        const std::string v = get(section, key);
        if (v == kMissing)
            die(code, "Property \"%s\" is not specified in the section \"%s\" in the input file \"%s\".", key, section,
                path.c_str());
        return v;
    }
};

Ini read_ini(const char *path) {
    Ini ini;
    ini.path = path;
    FILE *f = fopen(path, "rb");
    if (!f) die(1001, "The input file \"%s\" does not exist.", path);
    std::string text;
    char buf[65536];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
    fclose(f);
    std::string cur;
    bool have = false;
    size_t i = 0;
    while (i < text.size()) {
        size_t j = text.find('\n', i);
        if (j == std::string::npos) j = text.size();
        const std::string line = trim(text.substr(i, j - i));
        i = j + 1;
        if (line.empty() || line[0] == ';' || line[0] == '#') continue;
        if (line[0] == '[' && line.find(']') != std::string::npos) {
            cur = Ini::lower(trim(line.substr(1, line.find(']') - 1)));
            ini.sec[cur];
            have = true;
            continue;
        }
        const size_t eq = line.find('=');
        if (!have || eq == std::string::npos) continue;
        const std::string k = Ini::lower(trim(line.substr(0, eq)));
        std::string v = trim(line.substr(eq + 1));
        if (v.size() >= 2 && v.front() == '"' && v.back() == '"') v = v.substr(1, v.size() - 2);
        ini.sec[cur].emplace(k, v);
    }
    return ini;
}

// sscanf's "%f" / "%d" conversions of n blank-separated numbers: each takes
// the longest number at its position (floats rounded to binary32 through
// binary64); false when fewer than n convert
bool scan_numbers(const std::string &v, size_t n, bool integer, std::vector<double> &out) {
    out.clear();
    size_t p = 0;
    for (size_t k = 0; k < n; ++k) {
        while (p < v.size() && isspace((unsigned char)v[p])) ++p;
        const size_t s0 = p;
        if (p < v.size() && (v[p] == '+' || v[p] == '-')) ++p;
        const size_t d0 = p;
        while (p < v.size() && isdigit((unsigned char)v[p])) ++p;
        bool digits = p > d0;
        if (!integer) {
            if (p < v.size() && v[p] == '.') {
                ++p;
                const size_t f0 = p;
                while (p < v.size() && isdigit((unsigned char)v[p])) ++p;
                digits = digits || p > f0;
            }
            if (digits && p < v.size() && (v[p] == 'e' || v[p] == 'E')) {
                size_t q = p + 1;
                if (q < v.size() && (v[q] == '+' || v[q] == '-')) ++q;
                const size_t e0 = q;
                while (q < v.size() && isdigit((unsigned char)v[q])) ++q;
                if (q > e0) p = q;
            }
        }
        if (!digits) return false;
        const std::string t = v.substr(s0, p - s0);
        out.push_back(integer ? (double)strtoll(t.c_str(), nullptr, 10) : (double)(float)strtod(t.c_str(), nullptr));
    }
    return true;
}

std::vector<double> need_numbers(const Ini &ini, const std::string &v, size_t n, bool integer, int code,
                                 const char *what) {
    std::vector<double> out;
    if (!scan_numbers(v, n, integer, out))
        die(code, "Could not parse the %s \"%s\" in the input file \"%s\".", what, v.c_str(), ini.path.c_str());
    return out;
}

// mkdir -p
void make_dirs(const std::string &dir) {
    for (size_t k = 0; k <= dir.size(); ++k)
        if (k == dir.size() || (dir[k] == '/' && k > 0)) mkdir(dir.substr(0, k).c_str(), 0777);
}

uint32_t parse_fixes(const std::string &list, uint32_t &host, bool &q10) {
    uint32_t dev = 0;
    host = 0;
    q10 = false;
    size_t i = 0;
    while (i <= list.size()) {
        size_t j = list.find(',', i);
        if (j == std::string::npos) j = list.size();
        std::string f = Ini::lower(trim(list.substr(i, j - i)));
        i = j + 1;
        if (f.empty()) continue;
        if (f == "q1") dev |= GEGS_FIX_Q1;
        else if (f == "q7") dev |= GEGS_FIX_Q7;
        else if (f == "q8") dev |= GEGS_FIX_Q8;
        else if (f == "q4") host |= GEGS_HOST_FIX_Q4;
        else if (f == "q15") host |= GEGS_HOST_FIX_Q15;
        else if (f == "q16") host |= GEGS_HOST_FIX_Q16;
        else if (f == "q9") host |= GEGS_HOST_FIX_Q9;
        else if (f == "q10") q10 = true;                                   // the CT ramp, applied while reading
        else die(2001, "unknown fix \"%s\" (known here: Q1, Q4, Q7, Q8, Q9, Q10, Q15, Q16)", f.c_str());
    }
    return dev;
}

// The run input's transport part (CUDA_EGS.cu:936-1480 for the keys,
// init.cu:345-905 and :1377-1502 for the head and fields), as run.prepare
// restates it.
RunFile from_ini(const char *path, const std::string &media, const std::string &base_dir, const std::string &method,
                 uint64_t seed, uint32_t batches, const std::string &fixes, Log &log) {
    const Ini ini = read_ini(path);
    RunFile r;
    r.media = media;
    r.method = method;
    r.seed = seed;
    r.batches = batches;
    bool fix_q10 = false;
    r.device_fixes = parse_fixes(fixes, r.host_fixes, fix_q10);
    const std::string base = base_dir.empty() ? cwd() : base_dir;
    std::string prefix = ini.get("General", "output prefix");
    if (prefix == kMissing) {
        prefix = "GEGS";
        log("WARNING: No output prefix was specified. Default value is used.");
    }
    std::string outdir = ini.get("General", "output directory");
    if (outdir == kMissing || outdir.empty()) {
        outdir = ".";
        log("WARNING: No output directory was specified. Current Working Directory is used.");
    }
    for (const char *k : {"GPU number", "GPU id"}) {
        const std::string g = ini.get("General", k);
        std::vector<double> tmp;
        if (g != kMissing && !scan_numbers(g, 1, true, tmp))
            die(2004, "The given %s \"%s\" could not be parsed as a 4-byte integer.", k, g.c_str());
    }
    const std::string h = ini.req("General", "histories", 2013);
    std::vector<double> hv;
    if (!scan_numbers(h, 1, false, hv)) die(2014, "Could not parse the number of histories \"%s\".", h.c_str());
    const float hist = (float)hv[0];
    if (!(hist > 0)) die(2015, "The number of histories must be positive, but %g was specified.", (double)hist);
    r.histories_per_mu = (uint64_t)hist;
    const std::string data_dir = host_path(base, ini.req("Data", "data directory", 2017));
    const std::string egsinp = ini.req("Data", "egsinp file", 2011);
    ini.req("Data", "pegs file", 2012);
    ini.req("Data", "photon xsections", 2019);
    ini.req("Data", "atomic ff file", 2020);
    const std::string dmed = ini.req("Data", "default medium", 2013);
    const std::string nph = ini.req("Phase space", "phsp file number", 5013);
    std::vector<double> v;
    if (!scan_numbers(nph, 1, true, v)) die(5014, "Could not parse phsp file number \"%s\".", nph.c_str());
    r.phsp_files = (uint32_t)v[0];
    const std::string phsp = ini.req("Phase space", "phase space file", 2012);
    const std::string zs = ini.req("Phase space", "z position", 5013);
    if (!scan_numbers(zs, 1, false, v)) die(5014, "Could not parse the z position \"%s\".", zs.c_str());
    r.phsp_z = (float)v[0];
    const bool secjaws = ini.get("SecJaws", "module SecJaws") == "On";
    const std::string secjaws_file = ini.req("SecJaws", "secjaws file", 2011);
    const bool mlc_on = ini.get("MLC", "module MLC") == "On";
    const std::string mlcinfo = ini.req("MLC", "mlcinfo file", 2011);
    const bool block_on = ini.get("Block", "module Block") == "On";
    const bool wedge_on = ini.get("Wedge", "module Wedge") == "On";
    if (block_on) ini.req("Block", "block file", 2011);
    if (wedge_on) ini.req("Wedge", "wedge file", 2011);
    // the phantom: ctcreate from a DICOM series (CUDA_EGS.cu:1327-1383), or an egsphant file
    const bool use_ct = ini.get("CTcreate", "use ctcreate") == "Yes";
    std::string egsphant, ct_ramps, dicom_list;
    std::vector<double> ct_box, ct_voxel;
    if (use_ct) {
        ct_ramps = ini.req("CTcreate", "CT ramps file", 2012);
        dicom_list = ini.get("CTcreate", "dicom file list");
        if (dicom_list == kMissing) dicom_list = ini.req("CTcreate", "dicom directory", 2012);
        ct_box = need_numbers(ini, ini.req("CTcreate", "xyz boundaries", 2011), 6, false, 2011,
                              "boundaries of the subset of CT");
        ct_voxel = need_numbers(ini, ini.req("CTcreate", "3d voxel sizes", 2011), 3, false, 2011,
                                "voxel dimensions of CT");
    } else {
        egsphant = ini.req("Phantom", "egsphant file", 2011);
    }
    r.change_xyz = ini.get("Phantom", "change coordinate") != "No" ? 1 : 0;
    v = need_numbers(ini, ini.req("Phantom", "isocenter location", 2011), 3, false, 2011, "isocenter location");
    for (int k = 0; k < 3; ++k) r.iso[k] = (float)v[k];
    v = need_numbers(ini, ini.req("Phantom", "source to isocenter", 2011), 1, false, 2011, "source to isocenter");
    r.sid = (float)v[0];
    float dose_unit = 1.6e-8f;                                             // CUDA_EGS.h:536, MeV/g -> cGy
    if (ini.get("Calibration", "dose calibration") == "Yes") {
        const std::string hp = ini.req("Calibration", "histories per cGy", 5013);
        if (!scan_numbers(hp, 1, false, v)) die(5014, "Could not parse histories per cGy \"%s\".", hp.c_str());
        float du = dose_unit * (float)v[0];
        dose_unit = du / hist;                                             // CUDA_EGS.cu:1478, in float
    }
    r.dose_unit = dose_unit;
    const bool of_on = ini.get("Output factor", "use output factor") == "Yes";
    std::string of_file;
    if (of_on) of_file = ini.req("Output factor", "output factor file", 2012);
    if (ini.get("Interpolation", "do interpolate") == "Yes") {
        // CUDA_EGS.cu:1545-1560: the starting point moves by half a voxel
        const auto st = need_numbers(ini, ini.req("Interpolation", "starting point", 2011), 3, false, 2011,
                                     "starting point of interpolation");
        const auto dm = need_numbers(ini, ini.req("Interpolation", "xyz dimensions", 2011), 3, true, 2011,
                                     "xyz dimensions of interpolation");
        const auto sz = need_numbers(ini, ini.req("Interpolation", "3d voxel sizes", 2011), 3, false, 2011,
                                     "3d voxel sizes of interpolation");
        for (int k = 0; k < 3; ++k) {
            r.interp_sizes[k] = (float)sz[k];
            r.interp_dimen[k] = (uint32_t)dm[k];
            r.interp_start[k] = (float)st[k] + r.interp_sizes[k] / 2.0f;
        }
        r.interp = true;
    }
    if (ini.get("Dose output", "absolute dose") == "Yes") {
        const auto ps = need_numbers(ini, ini.req("Dose output", "dose position", 2011), 3, false, 2011,
                                     "dose position of absolute dose");
        const auto ms = need_numbers(ini, ini.req("Dose output", "dose measured", 2011), 1, false, 2011,
                                     "dose measured of absolute dose");
        for (int k = 0; k < 3; ++k) r.abs_pos[k] = (float)ps[k];
        r.abs_measured = (float)ms[0];
        r.abs_calibrated = ini.get("Calibration", "dose calibration") == "Yes" ? 1 : 0;
        r.absdose = true;
    }
    if (ini.get("Dose output", "relative dose") == "Yes") {
        // CUDA_EGS.cu:1600-1680: the starting point is relative to the isocentre
        const auto st = need_numbers(ini, ini.req("Dose output", "starting point", 2011), 3, false, 2011,
                                     "starting point of relative dose");
        const auto dm = need_numbers(ini, ini.req("Dose output", "xyz dimensions", 2011), 3, true, 2011,
                                     "xyz dimensions of relative dose");
        const auto sz = need_numbers(ini, ini.req("Dose output", "3d voxel sizes", 2011), 3, false, 2011,
                                     "3d voxel sizes of relative dose");
        for (int k = 0; k < 3; ++k) {
            r.rel_start[k] = (float)st[k] + r.iso[k];
            r.rel_dimen[k] = (uint32_t)dm[k];
            r.rel_sizes[k] = (float)sz[k];
            r.rel_iso[k] = r.iso[k];
        }
        const std::string at = ini.get("Dose output", "PTW729/PTW1000");
        r.rel_array = at == "PTW729" ? 1 : (at == "PTW1000" ? 2 : 0);
        const std::string nt = ini.get("Dose output", "norm to Max/Center");
        r.rel_norm = nt == "Max" ? 1 : (nt == "Center" ? 2 : 0);
        const auto gm = need_numbers(ini, ini.req("Dose output", "gamma index (mm,%)", 2011), 2, false, 2011,
                                     "gamma index paremeters");
        r.rel_glen = (float)gm[0] * 0.1f;
        r.rel_gdose = (float)gm[1] * 0.01f;
        r.rel_file = ini.req("Dose output", "measured data", 2012);     // resolved in the data directory below
        r.rel_measured = r.absdose ? r.abs_measured : 0.0f;
        r.abs_calibrated = ini.get("Calibration", "dose calibration") == "Yes" ? 1 : 0;
        r.reldose = true;
    }
    auto d = [&](const std::string &name) { return join(data_dir, slashes(name)); };
    if (r.reldose) r.rel_file = abspath(d(r.rel_file));
    const std::string out_dir = host_path(base, outdir);
    r.root = abspath(join(out_dir, prefix));
    if (use_ct) {
        // ctcreate writes <output root>.egsphant and the run reads the phantom back from it
        gegs_ct_ramp ramp[GEGS_CT_MAX_RAMP];
        uint32_t n_ramp = 0;
        int32_t ct_lower = 0;
        int rc = gegs_read_ct_ramps(d(ct_ramps).c_str(), ramp, &n_ramp, &ct_lower);
        if (rc) die(rc, "%s", gegs_ct_last_error());
        const float thick[3] = {(float)ct_voxel[0], (float)ct_voxel[1], (float)ct_voxel[2]};
        const float lo[3] = {(float)ct_box[0], (float)ct_box[2], (float)ct_box[4]};
        const float hi[3] = {(float)ct_box[1], (float)ct_box[3], (float)ct_box[5]};
        gegs_egsphant ct;
        rc = gegs_ctcreate(host_path(base, dicom_list).c_str(), thick, lo, hi, ramp, n_ramp, ct_lower, fix_q10 ? 1 : 0,
                           &ct);
        if (rc) die(rc, "%s", gegs_ct_last_error());
        make_dirs(out_dir);
        r.egsphant = abspath(join(out_dir, prefix) + ".egsphant");
        rc = gegs_egsphant_write_ctcreate(r.egsphant.c_str(), &ct);
        gegs_egsphant_free(&ct);
        if (rc) die(rc, "%s", gegs_ct_last_error());
    } else {
        r.egsphant = abspath(d(egsphant));
    }
    r.phsp_base = abspath(join(data_dir, slashes(phsp)));

    // head components, read whether or not the modules are on
    gegs_head &hd = r.head_mem;
    memset(&hd, 0, sizeof hd);
    char jaw_med[GEGS_NAME_LEN], mlc_med[GEGS_NAME_LEN], block_med[GEGS_NAME_LEN], wedge_med[GEGS_NAME_LEN];
    for (char *m : {jaw_med, mlc_med, block_med, wedge_med}) {
        memset(m, 0, GEGS_NAME_LEN);
        strncpy(m, dmed.c_str(), GEGS_NAME_LEN - 1);
    }
    if (exists(d(secjaws_file))) {
        const int rc = gegs_read_secjaws(d(secjaws_file).c_str(), &hd.jaws, jaw_med);
        if (rc != GEGS_OK && rc != GEGS_NOT_FOUND) die(rc, "%s", gegs_input_last_error());
        hd.has_jaws = rc == GEGS_OK;
        if (rc == GEGS_NOT_FOUND) strncpy(jaw_med, dmed.c_str(), GEGS_NAME_LEN - 1);
    }
    if (secjaws && !hd.has_jaws) die(2011, "no JAWS component SECJAWS in \"%s\"", d(secjaws_file).c_str());
    if (exists(d(egsinp))) {
        char med[GEGS_NAME_LEN];
        if (gegs_read_dynvmlc(d(egsinp).c_str(), &hd.mlc, med) == GEGS_OK) {
            hd.has_mlc = 1;
            memcpy(mlc_med, med, GEGS_NAME_LEN);
        } else {
            memset(&hd.mlc, 0, sizeof hd.mlc);
        }
    }
    if (mlc_on && !hd.has_mlc) die(2011, "no DYNVMLC component in \"%s\"", d(egsinp).c_str());
    float block_coord[10], wedge_rmax = 0, layers[24];
    uint32_t n_layers = 0;
    if (block_on) {
        const std::string bf = d(trim(ini.get("Block", "block file")));
        const int rc = gegs_read_block(bf.c_str(), block_coord, block_med);
        if (rc == GEGS_NOT_FOUND) die(2011, "no PYRAMIDS component BLOCK in the block file");
        if (rc) die(rc, "%s", gegs_input_last_error());
    }
    if (wedge_on) {
        const std::string wf = d(trim(ini.get("Wedge", "wedge file")));
        const int rc = gegs_read_wedge(wf.c_str(), &wedge_rmax, layers, &n_layers, wedge_med);
        if (rc == GEGS_NOT_FOUND) die(2011, "no PYRAMIDS component WEDGE in the wedge file");
        if (rc) die(rc, "%s", gegs_input_last_error());
    }
    if (block_on || wedge_on) {
        const int rc = gegs_build_block_wedge(block_on ? block_coord : nullptr, wedge_rmax, wedge_on ? layers : nullptr,
                                              wedge_on ? n_layers : 0, &hd.block_wedge);
        if (rc) die(rc, "%s", gegs_input_last_error());
        hd.has_block_wedge = 1;
    }
    const char *others[5] = {dmed.c_str(), jaw_med, mlc_med, block_med, wedge_med};
    for (int i = 0; i < 5; ++i) {
        memset(r.other[i], 0, GEGS_NAME_LEN);
        strncpy(r.other[i], others[i], GEGS_NAME_LEN - 1);
    }
    r.module_first = secjaws ? GEGS_M_SECJAWY : mlc_on ? GEGS_M_MLC : block_on ? GEGS_M_BLOCK
                   : wedge_on ? GEGS_M_WEDGE : GEGS_M_PHANTOM;

    // fields (read_mlcinp)
    const std::string mi = d(mlcinfo);
    if (!exists(mi)) die(2011, "Error: MLC info  file reading error! No simulation will be done!");
    uint32_t nf = 0;
    if (gegs_read_mlcinfo(mi.c_str(), nullptr, 0, &nf)) die(2011, "%s", gegs_input_last_error());
    std::vector<gegs_mlc_field> mf(nf ? nf : 1);
    if (gegs_read_mlcinfo(mi.c_str(), mf.data(), nf, &nf)) die(2011, "%s", gegs_input_last_error());
    const uint32_t kMaxFields = 20;                                        // CUDA_EGS.h:266
    if (nf > kMaxFields) {
        log("Warning! number of MLC fields is larger than the MAXFIELDS=%u", kMaxFields);
        nf = kMaxFields;
    }
    std::vector<float> sizes(64), factors(64);
    uint32_t nof = 0;
    if (of_on) {
        if (gegs_read_output_factors(d(of_file).c_str(), nullptr, nullptr, 0, &nof)) die(2011, "%s", gegs_input_last_error());
        sizes.resize(nof ? nof : 1);
        factors.resize(nof ? nof : 1);
        if (gegs_read_output_factors(d(of_file).c_str(), sizes.data(), factors.data(), nof, &nof))
            die(2011, "%s", gegs_input_last_error());
    }
    r.leaf_mem.assign((size_t)nf * 2 * GEGS_MLC_LEAVES, 0.0f);
    r.jaw_mem.assign((size_t)nf * 4, 0.0f);
    for (uint32_t i = 0; i < nf; ++i) {
        const gegs_mlc_field &f = mf[i];
        if (gegs_read_leaf_file(d(f.leaf_file).c_str(), r.leaf_mem.data() + (size_t)i * 2 * GEGS_MLC_LEAVES))
            die(2011, "%s", gegs_input_last_error());
        const float jp[4] = {(float)f.x1, (float)f.x2, (float)f.y1, (float)f.y2};
        memcpy(r.jaw_mem.data() + (size_t)i * 4, jp, sizeof jp);
        FieldSpec fs;
        fs.mu = (uint64_t)f.mu;
        fs.angle = (float)f.angle;
        // CUDA_EGS.cu:424: the angle in radians as binary32, its cos and sin
        // in binary64 rounded once
        const float a = (float)((double)fs.angle * 3.1415926 / 180.0);
        fs.cos_a = (float)std::cos((double)a);
        fs.sin_a = (float)std::sin((double)a);
        if (of_on) {
            float of = 0;
            const int rc = gegs_field_output_factor(f.x1, f.x2, f.y1, f.y2, sizes.data(), factors.data(), nof, &of);
            if (rc == GEGS_NOT_FOUND)
                die(2011, "field #%u: its equivalent square lies outside the output-factor table (the reference "
                          "leaves its factor unset)", i);
            fs.has_of = true;
            fs.output_factor = of;
        }
        r.fields.push_back(fs);
    }
    hd.n_sets = nf;
    hd.leaf_positions = hd.has_mlc ? r.leaf_mem.data() : nullptr;
    hd.jaw_positions = hd.has_jaws ? r.jaw_mem.data() : nullptr;
    r.head_in_memory = true;
    return r;
}

// the run file of a run (the format read_run_file reads and run.prepare writes)
void write_run_file(const RunFile &r, const char *path, const char *source) {
    const std::string stem = [&] {
        std::string p = abspath(path);
        const size_t slash = p.rfind('/'), dot = p.rfind('.');
        return (dot != std::string::npos && (slash == std::string::npos || dot > slash)) ? p.substr(0, dot) : p;
    }();
    const bool any_head = r.head_mem.has_mlc || r.head_mem.has_jaws || r.head_mem.has_block_wedge;
    const std::string head_file = stem + ".gegshead";
    if (any_head) {
        const int rc = gegs_head_file_write(head_file.c_str(), &r.head_mem);
        if (rc) die(rc, "%s", gegs_host_last_error());
    }
    FILE *f = fopen(path, "w");
    if (!f) die(2001, "Could not write the run file \"%s\".", path);
    auto g = [](float x) {
        char b[64];
        snprintf(b, sizeof b, "%.9g", (double)x);
        return std::string(b);
    };
    fprintf(f, "# gegs run file: %s prepared for gegs_amd/gegs_run\n", abspath(source).c_str());
    fprintf(f, "media file = %s\n", abspath(r.media).c_str());
    if (any_head) fprintf(f, "head file = %s\n", head_file.c_str());
    fprintf(f, "egsphant file = %s\n", r.egsphant.c_str());
    for (int i = 0; i < 5; ++i) fprintf(f, "region medium %d = %s\n", i + 1, r.other[i]);
    fprintf(f, "region flags = 0x%04x\n", r.flags);
    fprintf(f, "phsp base = %s\n", r.phsp_base.c_str());
    fprintf(f, "phsp files = %u\n", r.phsp_files);
    fprintf(f, "phsp z = %s\n", g(r.phsp_z).c_str());
    fprintf(f, "host fixes = %u\n", r.host_fixes);
    fprintf(f, "device fixes = %u\n", r.device_fixes);
    fprintf(f, "change xyz = %d\n", r.change_xyz);
    fprintf(f, "module first = %d\n", r.module_first);
    fprintf(f, "source to isocenter = %s\n", g(r.sid).c_str());
    fprintf(f, "isocenter = %s %s %s\n", g(r.iso[0]).c_str(), g(r.iso[1]).c_str(), g(r.iso[2]).c_str());
    fprintf(f, "histories per mu = %llu\n", (unsigned long long)r.histories_per_mu);
    fprintf(f, "dose unit = %s\n", g(r.dose_unit).c_str());
    fprintf(f, "output root = %s\n", r.root.c_str());
    fprintf(f, "seed = %llu\n", (unsigned long long)r.seed);
    fprintf(f, "batches = %u\n", r.batches);
    fprintf(f, "method = %s\n", r.method.c_str());
    if (r.interp)
        fprintf(f, "interpolation = %s %s %s %s %s %s %u %u %u\n", g(r.interp_start[0]).c_str(), g(r.interp_start[1]).c_str(),
                g(r.interp_start[2]).c_str(), g(r.interp_sizes[0]).c_str(), g(r.interp_sizes[1]).c_str(),
                g(r.interp_sizes[2]).c_str(), r.interp_dimen[0], r.interp_dimen[1], r.interp_dimen[2]);
    if (r.absdose)
        fprintf(f, "absolute dose = %s %s %s %d %s\n", g(r.abs_pos[0]).c_str(), g(r.abs_pos[1]).c_str(),
                g(r.abs_pos[2]).c_str(), r.abs_calibrated, g(r.abs_measured).c_str());
    if (r.reldose)
        fprintf(f, "relative dose = %s %s %s %s %s %s %u %u %u %s %s %s %d %d %s %s %d %s %s\n", g(r.rel_start[0]).c_str(),
                g(r.rel_start[1]).c_str(), g(r.rel_start[2]).c_str(), g(r.rel_sizes[0]).c_str(),
                g(r.rel_sizes[1]).c_str(), g(r.rel_sizes[2]).c_str(), r.rel_dimen[0], r.rel_dimen[1], r.rel_dimen[2],
                g(r.rel_iso[0]).c_str(), g(r.rel_iso[1]).c_str(), g(r.rel_iso[2]).c_str(), r.rel_norm, r.rel_array,
                g(r.rel_glen).c_str(), g(r.rel_gdose).c_str(), r.abs_calibrated, g(r.rel_measured).c_str(),
                r.rel_file.c_str());
    for (const FieldSpec &fs : r.fields)
        fprintf(f, "field = %llu %s %s %s %s\n", (unsigned long long)fs.mu, g(fs.angle).c_str(), g(fs.cos_a).c_str(),
                g(fs.sin_a).c_str(), fs.has_of ? g(fs.output_factor).c_str() : "-");
    fclose(f);
}

void check(gegs_ctx *ctx, int rc, const char *what) {
    if (rc) die(rc, "%s: %s", what, ctx ? gegs_last_error(ctx) : "");
}

void check_host(int rc, const char *what) {
    if (rc) die(rc, "%s: %s", what, gegs_host_last_error());
}

}  // namespace

int main(int argc, char **argv) {
    const char *usage = "usage: gegs_run -i <file.ini> --media <tables.gegsmed> [--base-dir D] [--method batched|plain] "
                        "[--seed S] [--batches B] [--fix Q1,Q4,...] [--dry-run <run file>] [--device N]\n"
                        "       gegs_run <run file> [--device N]";
    if (argc < 2) die(2001, "%s", usage);
    int device = 0;
    const char *ini_path = nullptr, *dry_run = nullptr;
    std::string media, base_dir, method = "batched", fixes;
    uint64_t seed = 1;
    uint32_t batches = 10;
    for (int i = 1; i < argc; ++i) {
        const bool more = i + 1 < argc;
        if (!strcmp(argv[i], "--device") && more) device = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-i") && more) ini_path = argv[++i];
        else if (!strcmp(argv[i], "--media") && more) media = argv[++i];
        else if (!strcmp(argv[i], "--base-dir") && more) base_dir = argv[++i];
        else if (!strcmp(argv[i], "--method") && more) method = argv[++i];
        else if (!strcmp(argv[i], "--seed") && more) seed = strtoull(argv[++i], nullptr, 10);
        else if (!strcmp(argv[i], "--batches") && more) batches = (uint32_t)strtoul(argv[++i], nullptr, 10);
        else if (!strcmp(argv[i], "--fix") && more) fixes = argv[++i];
        else if (!strcmp(argv[i], "--dry-run") && more) dry_run = argv[++i];
        else if (i > 1 || argv[i][0] == '-') die(2001, "unknown argument \"%s\"\n%s", argv[i], usage);
    }
    const auto t0 = std::chrono::steady_clock::now();
    Log log;
    RunFile rf;
    if (ini_path) {
        if (media.empty()) die(2001, "-i needs --media <tables.gegsmed> (python -m gegs_amd.hatch writes one)");
        if (method != "batched" && method != "plain") die(2001, "method must be batched or plain");
        log("GEGS (MI355X, C++ host) run input %s", ini_path);
        rf = from_ini(ini_path, media, base_dir, method, seed, batches, fixes, log);
        if (dry_run) {
            write_run_file(rf, dry_run, ini_path);
            printf("run file %s\n", dry_run);
            return 0;
        }
        make_dirs(rf.root.substr(0, rf.root.rfind('/')));
    } else {
        rf = read_run_file(argv[1]);
        log("GEGS (MI355X, C++ host) run file %s", argv[1]);
    }
    if (!log.open(rf.root + "_log.txt")) die(2001, "Could not open \"%s_log.txt\".", rf.root.c_str());

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
    if (rf.head_in_memory) head = rf.head_mem;
    else if (!rf.head.empty()) check_host(gegs_head_file_read(rf.head.c_str(), &head), "head file");
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
    std::vector<float> dose(nvox);
    check_host(gegs_write_dose(rf.root.c_str(), ph.nx, ph.ny, ph.nz, edep.data(), mass.data(), rf.dose_unit,
                               dose.data(), fname, sizeof fname),
               "dose file");
    log("%s", fname);
    // the other outputs of write_dose (output.c:873-918) on the host's voxel coordinates
    std::vector<float> xc(ph.nx + 1), yc(ph.ny + 1), zc(ph.nz + 1);
    gegs_host_bounds(ph.xb, ph.nx, xc.data());
    gegs_host_bounds(ph.yb, ph.ny, yc.data());
    gegs_host_bounds(ph.zb, ph.nz, zc.data());
    if (rf.interp) {
        std::vector<float> grid((size_t)rf.interp_dimen[0] * rf.interp_dimen[1] * rf.interp_dimen[2]);
        check_host(gegs_interp_dose(dose.data(), ph.nx, ph.ny, ph.nz, xc.data(), yc.data(), zc.data(), rf.interp_start,
                                    rf.interp_sizes, rf.interp_dimen, grid.data()),
                   "interpolated dose");
        char iname[4200];
        check_host(gegs_write_interp_dose(rf.root.c_str(), grid.data(), rf.interp_dimen, iname, sizeof iname),
                   "interpolated dose file");
        log("%s", iname);
    }
    auto abs_report = [&](const float *p, float meas) {
        float val, vox;
        uint32_t ijk[3];
        check_host(gegs_abs_dose(dose.data(), ph.nx, ph.ny, ph.nz, xc.data(), yc.data(), zc.data(), p,
                                 (rf.host_fixes & GEGS_HOST_FIX_Q9) != 0, &val, ijk, &vox),
                   "absolute dose");
        if (rf.abs_calibrated && meas == 0.0f) {
            log("Absolute dose at (%.2f, %.2f, %.2f) = %.1f cGy", p[0], p[1], p[2], val);
            log("   voxel dose of (i=%2u, j=%2u, k=%2u) = %.1f cGy", ijk[0], ijk[1], ijk[2], vox);
        } else if (rf.abs_calibrated) {
            const float d0 = std::fabs(val - meas) / meas, d1 = std::fabs(vox - meas) / meas;
            log("Absolute dose at (%.2f, %.2f, %.2f):", p[0], p[1], p[2]);
            log("  measured = %.1f cGy, calculated = %.1f cGy, difference = %.1f %%", meas, val, d0 * 100.0f);
            log("  voxel dose of index (%2u, %2u, %2u) = %.1f cGy, difference = %.1f %%", ijk[0], ijk[1], ijk[2], vox,
                d1 * 100.0f);
        } else {
            log("Absolute dose at (%.2f, %.2f, %.2f) = %.3E cGy", p[0], p[1], p[2], val);
            log("   voxel dose of (i=%2u, j=%2u, k=%2u) = %.3E cGy", ijk[0], ijk[1], ijk[2], vox);
        }
    };
    if (rf.absdose) abs_report(rf.abs_pos, rf.abs_measured);
    if (rf.reldose) {
        if (rf.rel_dimen[0] == 1 && rf.rel_dimen[1] == 1 && rf.rel_dimen[2] == 1) {
            abs_report(rf.rel_start, rf.rel_measured);
        } else {
            std::vector<float> calc((size_t)rf.rel_dimen[0] * rf.rel_dimen[1] * rf.rel_dimen[2]);
            check_host(gegs_reldose_points(dose.data(), ph.nx, ph.ny, ph.nz, xc.data(), yc.data(), zc.data(),
                                           rf.rel_start, rf.rel_sizes, rf.rel_dimen, calc.data()),
                       "relative dose points");
            char rname[4200], summary[512];
            check_host(gegs_write_relative_dose(rf.root.c_str(), calc.data(), rf.rel_dimen, rf.rel_file.c_str(),
                                                rf.rel_start, rf.rel_sizes, rf.rel_iso, rf.rel_norm, rf.rel_array,
                                                rf.rel_glen, rf.rel_gdose, rname, sizeof rname, summary, sizeof summary),
                       "relative dose");
            log("%s", rname);
            log("%s", summary);
        }
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    log("Total CPU/GPU . . . . . . . . %.2f ms (%.2f min)", secs * 1e3, secs / 60.0);
    printf("%llu histories in %.2f s -> %s\n", (unsigned long long)transported, secs, fname);
    if (!rf.head_in_memory) gegs_head_free(&head);
    gegs_egsphant_free(&ph);
    fclose(log.f);
    return 0;
}
