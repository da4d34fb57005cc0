// host_io.cpp -- the run's file formats on the host (declared in gegs_host.h).
//
// Host-only C++ (no HIP): built with the host compiler and linked into
// libgegs_hip.so, so that a C or C++ host reads a run's phantom, phase-space
// files and tables and writes its dose without Python.  Each reader follows
// the reference's own reader (cited per function) in what it accepts and in
// its error codes; the arithmetic that decides a result bit (voxel masses,
// relative densities, the dose file's binary32 values, the batch sizes) is
// the reference's, in the same order.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gegs_host.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

struct File {
    FILE *f = nullptr;
    explicit File(const char *path, const char *mode) : f(fopen(path, mode)) {}
    ~File() {
        if (f) fclose(f);
    }
    File(const File &) = delete;
    File &operator=(const File &) = delete;
};

// FNV-1a, 64 bit, over a byte block (the media-table file's checksum).
uint64_t fnv1a(const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

// The reference trims names with its trim(s, false): leading and trailing
// blanks, tabs, carriage returns and line feeds go (init.cu:1180).
std::string trim(const std::string &s) {
    const char *ws = " \t\r\n";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return std::string();
    const size_t b = s.find_last_not_of(ws);
    return s.substr(a, b - a + 1);
}

bool read_line(FILE *f, std::string &out) {
    out.clear();
    int c;
    bool any = false;
    while ((c = fgetc(f)) != EOF) {
        any = true;
        out.push_back((char)c);
        if (c == '\n') break;
    }
    return any;
}

int find_medium(const char names[GEGS_MXMED][GEGS_NAME_LEN], int32_t nmed, const char *name) {
    for (int32_t m = 0; m < nmed; ++m)
        if (strncmp(names[m], name, GEGS_NAME_LEN) == 0) return m;
    return -1;
}

const uint64_t kMediaMagic = 0x3130544d53474547ull;   // "GEGSMT01"
const uint64_t kHeadMagic = 0x3130444853474547ull;    // "GEGSHD01"
const uint32_t kLeaves = 2 * GEGS_MLC_LEAVES;

// The reference's launch constants that fix a field's batch size
// (CUDA_EGS.h:111-138; init.cu:31: the photon stack is 1.5 batches).
const uint32_t kPhspStackSize = 50u * 5u * 15u * 4u * 32u;
const uint32_t kMaxStackSize = 10000000u;

}  // namespace

struct gegs_phsp {
    std::vector<std::string> header_paths, data_paths;
    std::vector<uint64_t> counts;
    uint64_t total = 0;
    bool fix_q15 = false;
};

extern "C" {

const char *gegs_host_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------------------------ media-table file
int gegs_media_file_write(const char *path, const gegs_media_tables *t, const char *names) {
    if (!path || !t || !names) return fail(GEGS_ERR_ARG, "gegs_media_file_write: null argument");
    if (t->nmed < 1 || t->nmed > GEGS_MXMED)
        return fail(GEGS_ERR_MEDIA, "gegs_media_file_write: nmed = %d outside 1..%d", t->nmed, GEGS_MXMED);
    File f(path, "wb");
    if (!f.f) return fail(GEGS_ERR_ARG, "Could not open \"%s\" for writing: %s", path, strerror(errno));
    const uint64_t size = sizeof(gegs_media_tables);
    const uint32_t nmed = (uint32_t)t->nmed;
    const uint64_t sum = fnv1a(t, sizeof *t);
    bool ok = fwrite(&kMediaMagic, 8, 1, f.f) == 1 && fwrite(&size, 8, 1, f.f) == 1 && fwrite(&nmed, 4, 1, f.f) == 1;
    for (uint32_t m = 0; ok && m < nmed; ++m) {
        char nm[GEGS_NAME_LEN] = {0};
        strncpy(nm, names + (size_t)m * GEGS_NAME_LEN, GEGS_NAME_LEN - 1);
        ok = fwrite(nm, GEGS_NAME_LEN, 1, f.f) == 1;
    }
    ok = ok && fwrite(t, sizeof *t, 1, f.f) == 1 && fwrite(&sum, 8, 1, f.f) == 1;
    if (!ok) return fail(GEGS_ERR_ARG, "Could not write \"%s\"", path);
    return GEGS_OK;
}

int gegs_media_file_read(const char *path, gegs_media_tables *t, char names[GEGS_MXMED][GEGS_NAME_LEN]) {
    if (!path || !t || !names) return fail(GEGS_ERR_ARG, "gegs_media_file_read: null argument");
    File f(path, "rb");
    if (!f.f) return fail(GEGS_ERR_MEDIA, "Could not open the media-table file \"%s\".", path);
    uint64_t magic = 0, size = 0, sum = 0;
    uint32_t nmed = 0;
    if (fread(&magic, 8, 1, f.f) != 1 || magic != kMediaMagic)
        return fail(GEGS_ERR_MEDIA, "\"%s\" is not a media-table file", path);
    if (fread(&size, 8, 1, f.f) != 1 || size != sizeof(gegs_media_tables))
        return fail(GEGS_ERR_MEDIA, "\"%s\" holds a table block of %llu bytes, this library's is %zu", path,
                    (unsigned long long)size, sizeof(gegs_media_tables));
    if (fread(&nmed, 4, 1, f.f) != 1 || nmed < 1 || nmed > GEGS_MXMED)
        return fail(GEGS_ERR_MEDIA, "\"%s\": medium count outside 1..%d", path, GEGS_MXMED);
    memset(names, 0, (size_t)GEGS_MXMED * GEGS_NAME_LEN);
    for (uint32_t m = 0; m < nmed; ++m) {
        if (fread(names[m], GEGS_NAME_LEN, 1, f.f) != 1) return fail(GEGS_ERR_MEDIA, "\"%s\" is truncated", path);
        names[m][GEGS_NAME_LEN - 1] = 0;
    }
    if (fread(t, sizeof *t, 1, f.f) != 1 || fread(&sum, 8, 1, f.f) != 1)
        return fail(GEGS_ERR_MEDIA, "\"%s\" is truncated", path);
    if (fnv1a(t, sizeof *t) != sum) return fail(GEGS_ERR_MEDIA, "\"%s\": checksum mismatch", path);
    if (t->nmed != (int32_t)nmed) return fail(GEGS_ERR_MEDIA, "\"%s\": the names and the tables disagree on nmed", path);
    return GEGS_OK;
}

// ------------------------------------------------------------------ egsphant
// init_phantom_and_MLC's reader (init.cu:1164-1251): fscanf-style tokens, so
// media digits and numbers may run across line breaks; after the last digit
// (density) of a z slice the rest of that line must be blank.
int gegs_egsphant_read(const char *path, gegs_egsphant *out) {
    if (!path || !out) return fail(GEGS_ERR_ARG, "gegs_egsphant_read: null argument");
    memset(out, 0, sizeof *out);
    File f(path, "r");
    if (!f.f) return fail(8001, "Could not open the phantom file \"%s\".", path);
    int nmed = 0;
    if (fscanf(f.f, "%d", &nmed) != 1 || nmed < 1)
        return fail(8002, "Could not read the number of media form the phantom file \"%s\".", path);
    // "%d\n" also skips the white space before the first name
    int c;
    while ((c = fgetc(f.f)) != EOF && isspace(c)) {
    }
    if (c != EOF) ungetc(c, f.f);
    std::vector<std::string> names;
    for (int i = 0; i < nmed; ++i) {
        std::string line;
        if (!read_line(f.f, line))
            return fail(8003, "Could not read medium name (i = %d) in the phantom file \"%s\".", i, path);
        line = trim(line);
        if (line.empty()) return fail(8004, "Medium name cannot be empty (i = %d) in the phantom file \"%s\".", i, path);
        if (line.size() >= GEGS_NAME_LEN)
            return fail(8004, "Medium name \"%s\" longer than %d characters in \"%s\".", line.c_str(), GEGS_NAME_LEN - 1,
                        path);
        names.push_back(line);
    }
    std::string dummy;
    read_line(f.f, dummy);   // the ESTEPE line
    int nx, ny, nz;
    if (fscanf(f.f, "%d %d %d", &nx, &ny, &nz) != 3 || nx < 1 || ny < 1 || nz < 1)
        return fail(8006, "Could not read the voxel numbers in the phantom file \"%s\".", path);
    const size_t nvox = (size_t)nx * ny * nz;
    std::vector<float> xb(nx + 1), yb(ny + 1), zb(nz + 1);
    for (int i = 0; i <= nx; ++i)
        if (fscanf(f.f, "%f", &xb[i]) != 1)
            return fail(8007, "Could not read x coordinate of voxel boundary (i = %d) in the phantom file \"%s\".", i,
                        path);
    for (int i = 0; i <= ny; ++i)
        if (fscanf(f.f, "%f", &yb[i]) != 1)
            return fail(8008, "Could not read y coordinate of voxel boundary (i = %d) in the phantom file \"%s\".", i,
                        path);
    for (int i = 0; i <= nz; ++i)
        if (fscanf(f.f, "%f", &zb[i]) != 1)
            return fail(8009, "Could not read z coordinate of voxel boundary (i = %d) in the phantom file \"%s\".", i,
                        path);
    std::vector<int32_t> media(nvox);
    std::vector<float> dens(nvox);
    for (int z = 0; z < nz; ++z) {
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                int v;
                if (fscanf(f.f, "%1d", &v) != 1)
                    return fail(8010, "Could not read media index (x = %d, y = %d, z = %d) in the phantom file \"%s\".",
                                x, y, z, path);
                if (v < 0 || v > nmed)
                    return fail(8010, "Media index %d outside 0..%d (x = %d, y = %d, z = %d) in \"%s\".", v, nmed, x, y,
                                z, path);
                media[(size_t)x + (size_t)y * nx + (size_t)z * nx * ny] = v;
            }
        std::string rest;
        read_line(f.f, rest);
        if (!trim(rest).empty())
            return fail(8011, "Expected empty line but found \"%s\" (z = %d) in the phantom file \"%s\".",
                        trim(rest).c_str(), z, path);
    }
    for (int z = 0; z < nz; ++z) {
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                if (fscanf(f.f, "%f", &dens[(size_t)x + (size_t)y * nx + (size_t)z * nx * ny]) != 1)
                    return fail(8012, "Could not read density (x = %d, y = %d, z = %d) in the phantom file \"%s\".", x, y,
                                z, path);
        std::string rest;
        read_line(f.f, rest);
        if (!trim(rest).empty())
            return fail(8013, "Expected empty line but found \"%s\" (z = %d) in the phantom file \"%s\".",
                        trim(rest).c_str(), z, path);
    }
    out->nx = (uint32_t)nx;
    out->ny = (uint32_t)ny;
    out->nz = (uint32_t)nz;
    out->nmed = (uint32_t)nmed;
    out->names = static_cast<char (*)[GEGS_NAME_LEN]>(calloc((size_t)nmed, GEGS_NAME_LEN));
    out->xb = static_cast<float *>(malloc(sizeof(float) * (nx + 1)));
    out->yb = static_cast<float *>(malloc(sizeof(float) * (ny + 1)));
    out->zb = static_cast<float *>(malloc(sizeof(float) * (nz + 1)));
    out->media = static_cast<int32_t *>(malloc(sizeof(int32_t) * nvox));
    out->density = static_cast<float *>(malloc(sizeof(float) * nvox));
    if (!out->names || !out->xb || !out->yb || !out->zb || !out->media || !out->density) {
        gegs_egsphant_free(out);
        return fail(GEGS_ERR_ARG, "out of host memory reading \"%s\"", path);
    }
    for (int i = 0; i < nmed; ++i) memcpy(out->names[i], names[i].c_str(), names[i].size() + 1);
    memcpy(out->xb, xb.data(), sizeof(float) * xb.size());
    memcpy(out->yb, yb.data(), sizeof(float) * yb.size());
    memcpy(out->zb, zb.data(), sizeof(float) * zb.size());
    memcpy(out->media, media.data(), sizeof(int32_t) * nvox);
    memcpy(out->density, dens.data(), sizeof(float) * nvox);
    return GEGS_OK;
}

void gegs_egsphant_free(gegs_egsphant *ph) {
    if (!ph) return;
    free(ph->names);
    free(ph->xb);
    free(ph->yb);
    free(ph->zb);
    free(ph->media);
    free(ph->density);
    memset(ph, 0, sizeof *ph);
}

int gegs_voxel_mass(const gegs_egsphant *ph, float *mass) {
    if (!ph || !ph->xb || !mass) return fail(GEGS_ERR_ARG, "gegs_voxel_mass: null argument");
    const uint32_t nx = ph->nx, ny = ph->ny, nz = ph->nz;
    for (uint32_t z = 0; z < nz; ++z)
        for (uint32_t y = 0; y < ny; ++y)
            for (uint32_t x = 0; x < nx; ++x) {
                const size_t i = (size_t)x + (size_t)y * nx + (size_t)z * nx * ny;
                float vol = (ph->xb[x + 1] - ph->xb[x]) * (ph->yb[y + 1] - ph->yb[y]);
                vol = vol * (ph->zb[z + 1] - ph->zb[z]);
                if (vol < 0) vol = -vol;
                float m = vol * ph->density[i];
                if (m == 0) m = 1;
                mass[i] = m;
            }
    return GEGS_OK;
}

int gegs_region_table(const gegs_egsphant *ph, const gegs_media_tables *t, const char names[GEGS_MXMED][GEGS_NAME_LEN],
                      const char other_media[5][GEGS_NAME_LEN], uint16_t flags, gegs_region *out) {
    if (!ph || !ph->media || !t || !names || !other_media || !out)
        return fail(GEGS_ERR_ARG, "gegs_region_table: null argument");
    out[0] = gegs_region{(uint16_t)GEGS_VACUUM, 0, 0.0f, 0.0f, 0.0f};
    for (int i = 1; i <= 5; ++i) {
        const int m = find_medium(names, t->nmed, other_media[i - 1]);
        if (m < 0) return fail(GEGS_ERR_MEDIA, "medium \"%.63s\" of region %d is not in the media tables", other_media[i - 1], i);
        out[i] = gegs_region{(uint16_t)m, flags, 1.0f, t->ap[m], t->ae[m]};
    }
    std::vector<int> lut(ph->nmed + 1, -1);
    for (uint32_t k = 0; k < ph->nmed; ++k) {
        lut[k + 1] = find_medium(names, t->nmed, ph->names[k]);
        if (lut[k + 1] < 0) return fail(GEGS_ERR_MEDIA, "phantom medium \"%s\" is not in the media tables", ph->names[k]);
    }
    const size_t nvox = (size_t)ph->nx * ph->ny * ph->nz;
    for (size_t v = 0; v < nvox; ++v) {
        const int32_t mi = ph->media[v];
        gegs_region &r = out[6 + v];
        if (mi == 0) {
            r = gegs_region{(uint16_t)GEGS_VACUUM, 0, 0.0f, 0.0f, 0.0f};
            continue;
        }
        const int m = lut[mi];
        const float d = ph->density[v];
        r.med = (uint16_t)m;
        r.flags = flags;
        r.rhof = d == 0.0f ? 1.0f : d / t->rho[m];
        r.pcut = t->ap[m];
        r.ecut = t->ae[m];
    }
    return GEGS_OK;
}

// ------------------------------------------------------------------ egsphsp
// read_egsphsp_header (read_egsphsp.cpp:60-165): the record count NPPHSP of
// every file; MODE0 records of 28 bytes after a 28-byte header.
int gegs_phsp_open(const char *base, uint32_t n_files, uint32_t fixes, gegs_phsp **out) {
    if (!base || !out) return fail(GEGS_ERR_ARG, "gegs_phsp_open: null argument");
    *out = nullptr;
    gegs_phsp *p = new gegs_phsp;
    p->fix_q15 = (fixes & GEGS_HOST_FIX_Q15) != 0;
    if (n_files > 1) {
        for (uint32_t i = 0; i < n_files; ++i) p->header_paths.push_back(std::string(base) + "_p" + std::to_string(i) + ".egsphsp1");
        p->data_paths = p->header_paths;
    } else {
        p->header_paths.push_back(std::string(base) + ".egsphsp1");
        p->data_paths.push_back(std::string(base) + ((fixes & GEGS_HOST_FIX_Q4) ? ".egsphsp1" : "_p0.egsphsp1"));
    }
    for (const std::string &hp : p->header_paths) {
        File f(hp.c_str(), "rb");
        unsigned char hdr[28];
        if (!f.f) {
            delete p;
            return fail(GEGS_ERR_ARG, "Could not open the egsphsp file \"%s\".", hp.c_str());
        }
        if (fread(hdr, 1, 28, f.f) != 28) {
            delete p;
            return fail(GEGS_ERR_ARG, "The egsphsp file \"%s\" has no complete header.", hp.c_str());
        }
        uint32_t n;
        memcpy(&n, hdr + 8, 4);
        p->counts.push_back(n);
        p->total += n;
    }
    *out = p;
    return GEGS_OK;
}

uint64_t gegs_phsp_total(const gegs_phsp *p) { return p ? p->total : 0; }

const char *gegs_phsp_header_path(const gegs_phsp *p, uint32_t i) {
    return p && i < p->header_paths.size() ? p->header_paths[i].c_str() : nullptr;
}

static int phsp_records(const gegs_phsp *p, size_t i, uint64_t start, uint64_t count, unsigned char *dst) {
    const std::string &path = p->data_paths[i];
    File f(path.c_str(), "rb");
    if (!f.f)
        return fail(GEGS_ERR_ARG, "Could not open the egsphsp file \"%s\"%s.", path.c_str(),
                    p->header_paths.size() <= 1 ? " (the reference reads the records of a single phase-space file from "
                                                  "<name>_p0.egsphsp1, quirk Q4; GEGS_HOST_FIX_Q4 reads <name>.egsphsp1)"
                                                : "");
    if (fseeko(f.f, (off_t)(28 + 28 * start), SEEK_SET) != 0 || fread(dst, 28, count, f.f) != count)
        return fail(GEGS_ERR_ARG, "The egsphsp file \"%s\" holds fewer than %llu records from record %llu.", path.c_str(),
                    (unsigned long long)count, (unsigned long long)start);
    return GEGS_OK;
}

// read_egsphsp_data (read_egsphsp.cpp:169-232) and the decoding of
// load_stack_from_phspfile (CUDA_EGS.cu:40-106).
int gegs_phsp_read(gegs_phsp *p, uint64_t istart, uint64_t length, float z, gegs_source *out) {
    if (!p || (!out && length)) return fail(GEGS_ERR_ARG, "gegs_phsp_read: null argument");
    if (istart >= p->total || p->total - istart < length) return GEGS_PHSP_REFUSED;
    size_t i = 0;
    while (i < p->counts.size() && istart >= p->counts[i]) istart -= p->counts[i++];
    if (i >= p->counts.size()) return GEGS_PHSP_REFUSED;
    std::vector<unsigned char> raw((size_t)length * 28);
    const uint64_t remain = p->counts[i] - istart;
    int rc;
    if (remain < length) {
        const size_t second = p->fix_q15 ? i + 1 : i;
        rc = phsp_records(p, i, istart, remain, raw.data());
        if (!rc) rc = phsp_records(p, second, 0, length - remain, raw.data() + remain * 28);
    } else {
        rc = phsp_records(p, i, istart, length, raw.data());
    }
    if (rc) return rc;
    for (uint64_t k = 0; k < length; ++k) {
        const unsigned char *r = raw.data() + k * 28;
        uint32_t latch;
        float e, x, y, u, v, wt;
        memcpy(&latch, r, 4);
        memcpy(&e, r + 4, 4);
        memcpy(&x, r + 8, 4);
        memcpy(&y, r + 12, 4);
        memcpy(&u, r + 16, 4);
        memcpy(&v, r + 20, 4);
        memcpy(&wt, r + 24, 4);
        const int32_t q = (latch & 0x40000000u) ? -1 : (latch & 0x20000000u) ? 1 : 0;
        float arg = 1.0f - u * u;
        arg = arg - v * v;
        float w = std::sqrt(arg > 0.0f ? arg : 0.0f);
        if (wt < 0) w = -w;
        gegs_source &s = out[k];
        s.e = std::fabs(e);
        s.x = x;
        s.y = y;
        s.z = z;
        s.u = u;
        s.v = v;
        s.w = w;
        s.wt = std::fabs(wt);
        s.charge = q;
    }
    return GEGS_OK;
}

void gegs_phsp_close(gegs_phsp *p) { delete p; }

// ------------------------------------------------------------------ batches of a field
int gegs_field_schedule(uint64_t histories, uint64_t total, uint32_t fixes, gegs_schedule *s, uint64_t *loads,
                        uint64_t max_loads) {
    if (!s) return fail(GEGS_ERR_ARG, "gegs_field_schedule: null argument");
    memset(s, 0, sizeof *s);
    if (histories == 0 || total == 0 || (histories - 1) / total > 255)
        return fail(GEGS_ERR_ARG, "The repeat number is not acceptable, try another number of hitories or phase space file");
    const uint64_t n_repeat = (histories - 1) / total;
    const uint32_t nr = (uint32_t)((n_repeat + 1) & 0xFFu);
    if (nr == 0)
        return fail(GEGS_ERR_ARG, "nRepeat = 255: the reference's unsigned-char repeat count wraps to 0 and its particle "
                                  "stacks are allocated empty (init.cu:33); use fewer histories per record");
    // init_stack (init.cu:29-41): the photon stack holds 1.5 batches
    uint32_t bs = kPhspStackSize;
    const float photon_factor = 1.5f;
    const float need = (float)(kPhspStackSize * nr) * photon_factor;
    if (need >= (float)kMaxStackSize) bs = (uint32_t)((float)kMaxStackSize / ((float)nr * photon_factor));
    const uint64_t r = n_repeat + 1;
    const uint64_t n_batch = (histories - 1) / ((uint64_t)bs * r);
    const uint64_t last = (histories - n_batch * ((uint64_t)bs * r)) / r + 1;
    const bool fix_q16 = (fixes & GEGS_HOST_FIX_Q16) != 0;
    s->histories = histories;
    s->n_repeat = (uint32_t)n_repeat;
    s->batch_size = bs;
    s->n_batch = n_batch;
    s->last_batch = last;
    for (uint64_t b = 0; b <= n_batch; ++b) {
        const uint64_t start = b * bs;
        uint64_t n = b < n_batch ? bs : last;
        if (fix_q16 && b == n_batch) n = (histories - n_batch * bs * r + r - 1) / r;
        if (start + n <= total) {
            if (loads && s->n_loads < max_loads) {
                loads[2 * s->n_loads] = start;
                loads[2 * s->n_loads + 1] = n;
            }
            s->n_loads++;
            s->records += n;
        } else {
            s->n_dropped++;
            s->dropped_records += n;
        }
    }
    if (fix_q16 && s->n_dropped == 0) {
        const uint64_t full = histories / r;
        for (uint64_t q = 0; q < r; ++q) s->repeat_counts[q] = full + (q < histories - full * r ? 1 : 0);
    } else {
        for (uint64_t q = 0; q < r; ++q) s->repeat_counts[q] = s->records;
    }
    return GEGS_OK;
}

// ------------------------------------------------------------------ dose file
int gegs_write_dose(const char *root, uint32_t nx, uint32_t ny, uint32_t nz, const double *edep, const float *mass,
                    float dose_unit, float *dose_out, char *fname, size_t fname_len) {
    if (!root || !edep || !mass) return fail(GEGS_ERR_ARG, "gegs_write_dose: null argument");
    char name[4096];
    snprintf(name, sizeof name, "%s_DoseResults_%03ux%03ux%03u.bin", root, nx, ny, nz);
    const size_t n = (size_t)nx * ny * nz;
    std::vector<float> d(n);
    for (size_t i = 0; i < n; ++i) {
        const float q = (float)edep[i] / mass[i];
        d[i] = q * dose_unit;
    }
    File f(name, "wb");
    if (!f.f || fwrite(d.data(), sizeof(float), n, f.f) != n)
        return fail(GEGS_ERR_ARG, "Could not write the dose file \"%s\"", name);
    if (dose_out) memcpy(dose_out, d.data(), n * sizeof(float));
    if (fname && fname_len) {
        strncpy(fname, name, fname_len - 1);
        fname[fname_len - 1] = 0;
    }
    return GEGS_OK;
}

// ------------------------------------------------------------------ head file
int gegs_head_file_write(const char *path, const gegs_head *h) {
    if (!path || !h) return fail(GEGS_ERR_ARG, "gegs_head_file_write: null argument");
    if ((h->has_mlc && !h->leaf_positions) || (h->has_jaws && !h->jaw_positions))
        return fail(GEGS_ERR_ARG, "gegs_head_file_write: aperture sets missing");
    File f(path, "wb");
    if (!f.f) return fail(GEGS_ERR_ARG, "Could not open \"%s\" for writing: %s", path, strerror(errno));
    const uint64_t magic = kHeadMagic;
    const uint32_t hdr[8] = {h->has_mlc, h->has_jaws, h->has_block_wedge, h->n_sets, (uint32_t)sizeof(gegs_mlc),
                             (uint32_t)sizeof(gegs_jaws), (uint32_t)sizeof(gegs_block_wedge), 0};
    bool ok = fwrite(&magic, 8, 1, f.f) == 1 && fwrite(hdr, sizeof hdr, 1, f.f) == 1 &&
              fwrite(&h->mlc, sizeof h->mlc, 1, f.f) == 1 && fwrite(&h->jaws, sizeof h->jaws, 1, f.f) == 1 &&
              fwrite(&h->block_wedge, sizeof h->block_wedge, 1, f.f) == 1;
    if (ok && h->has_mlc) ok = fwrite(h->leaf_positions, sizeof(float) * kLeaves, h->n_sets, f.f) == h->n_sets;
    if (ok && h->has_jaws) ok = fwrite(h->jaw_positions, sizeof(float) * 4, h->n_sets, f.f) == h->n_sets;
    if (!ok) return fail(GEGS_ERR_ARG, "Could not write \"%s\"", path);
    return GEGS_OK;
}

int gegs_head_file_read(const char *path, gegs_head *h) {
    if (!path || !h) return fail(GEGS_ERR_ARG, "gegs_head_file_read: null argument");
    memset(h, 0, sizeof *h);
    File f(path, "rb");
    if (!f.f) return fail(GEGS_ERR_ARG, "Could not open the head file \"%s\".", path);
    uint64_t magic = 0;
    uint32_t hdr[8];
    if (fread(&magic, 8, 1, f.f) != 1 || magic != kHeadMagic)
        return fail(GEGS_ERR_ARG, "\"%s\" is not a head file", path);
    if (fread(hdr, sizeof hdr, 1, f.f) != 1 || hdr[4] != sizeof(gegs_mlc) || hdr[5] != sizeof(gegs_jaws) ||
        hdr[6] != sizeof(gegs_block_wedge))
        return fail(GEGS_ERR_ARG, "\"%s\" was written for other structure sizes", path);
    h->has_mlc = hdr[0];
    h->has_jaws = hdr[1];
    h->has_block_wedge = hdr[2];
    h->n_sets = hdr[3];
    if (fread(&h->mlc, sizeof h->mlc, 1, f.f) != 1 || fread(&h->jaws, sizeof h->jaws, 1, f.f) != 1 ||
        fread(&h->block_wedge, sizeof h->block_wedge, 1, f.f) != 1)
        return fail(GEGS_ERR_ARG, "\"%s\" is truncated", path);
    if (h->has_mlc) {
        h->leaf_positions = static_cast<float *>(malloc(sizeof(float) * kLeaves * (h->n_sets ? h->n_sets : 1)));
        if (!h->leaf_positions || fread(h->leaf_positions, sizeof(float) * kLeaves, h->n_sets, f.f) != h->n_sets) {
            gegs_head_free(h);
            return fail(GEGS_ERR_ARG, "\"%s\": leaf positions truncated", path);
        }
    }
    if (h->has_jaws) {
        h->jaw_positions = static_cast<float *>(malloc(sizeof(float) * 4 * (h->n_sets ? h->n_sets : 1)));
        if (!h->jaw_positions || fread(h->jaw_positions, sizeof(float) * 4, h->n_sets, f.f) != h->n_sets) {
            gegs_head_free(h);
            return fail(GEGS_ERR_ARG, "\"%s\": jaw positions truncated", path);
        }
    }
    return GEGS_OK;
}

void gegs_head_free(gegs_head *h) {
    if (!h) return;
    free(h->leaf_positions);
    free(h->jaw_positions);
    h->leaf_positions = h->jaw_positions = nullptr;
}

}  // extern "C"

// ------------------------------------------------------------------ dose outputs of write_dose
namespace {

// one pass of interp_dose (output.c:222-312): values v[.., k] placed at
// b[k] linearly interpolated onto the coordinates c along the last axis, in
// binary32; points outside (b[0], b[n - 1]) stay 0
void axis_interp(const float *v, size_t outer, uint32_t n, const float *b, const float *c, uint32_t m, float *out) {
    for (size_t o = 0; o < outer; ++o) {
        const float *vo = v + o * n;
        float *w = out + o * m;
        for (uint32_t q = 0; q < m; ++q) {
            const float t = c[q];
            w[q] = 0.0f;
            if (t <= b[0] || t >= b[n - 1]) continue;
            for (uint32_t k = 0; k + 1 < n; ++k)
                if (t > b[k] && t <= b[k + 1]) {
                    float d = (t - b[k]) * (vo[k + 1] - vo[k]);
                    d = d / (b[k + 1] - b[k]);
                    w[q] = vo[k] + d;
                    break;
                }
        }
    }
}

uint32_t nearest_cell(const float *b, uint32_t n, float p) {
    uint32_t i = 1;
    while (i < n - 2) {
        if (p > b[i] && p <= b[i + 1]) break;
        ++i;
    }
    if (i == n - 2) i = p <= b[1] ? 0 : n - 2;
    return i;
}

}  // namespace

extern "C" {

int gegs_host_bounds(const float *b, uint32_t n, float *out) {
    if (!b || !out || n < 1) return fail(GEGS_ERR_ARG, "gegs_host_bounds: bad argument");
    for (uint32_t k = 0; k < n; ++k) out[k] = (b[k] + b[k + 1]) / 2.0f;
    out[n] = b[n];
    return GEGS_OK;
}

int gegs_interp_dose(const float *dose, uint32_t nx, uint32_t ny, uint32_t nz, const float *xc, const float *yc,
                     const float *zc, const float start[3], const float sizes[3], const uint32_t dimen[3], float *out) {
    if (!dose || !xc || !yc || !zc || !start || !sizes || !dimen || !out)
        return fail(GEGS_ERR_ARG, "gegs_interp_dose: null argument");
    if (nx < 2 || ny < 2 || nz < 2) return fail(GEGS_ERR_ARG, "gegs_interp_dose: at least 2 voxels per axis");
    std::vector<float> pts[3];
    for (int a = 0; a < 3; ++a) {
        pts[a].resize(dimen[a]);
        for (uint32_t i = 0; i < dimen[a]; ++i) pts[a][i] = start[a] + sizes[a] * (float)i;
    }
    const uint32_t dx = dimen[0], dy = dimen[1], dz = dimen[2];
    // z pass over every (y, x) column, then y, then x, as the reference does
    std::vector<float> col((size_t)ny * nx * nz), vz((size_t)ny * nx * dz);
    for (uint32_t j = 0; j < ny; ++j)
        for (uint32_t i = 0; i < nx; ++i)
            for (uint32_t k = 0; k < nz; ++k) col[((size_t)j * nx + i) * nz + k] = dose[((size_t)k * ny + j) * nx + i];
    axis_interp(col.data(), (size_t)ny * nx, nz, zc, pts[2].data(), dz, vz.data());      // [ny][nx][dz]
    std::vector<float> cy((size_t)nx * dz * ny), vy((size_t)nx * dz * dy);
    for (uint32_t j = 0; j < ny; ++j)
        for (uint32_t i = 0; i < nx; ++i)
            for (uint32_t k = 0; k < dz; ++k) cy[((size_t)i * dz + k) * ny + j] = vz[((size_t)j * nx + i) * dz + k];
    axis_interp(cy.data(), (size_t)nx * dz, ny, yc, pts[1].data(), dy, vy.data());       // [nx][dz][dy]
    std::vector<float> cx((size_t)dz * dy * nx);
    for (uint32_t i = 0; i < nx; ++i)
        for (uint32_t k = 0; k < dz; ++k)
            for (uint32_t j = 0; j < dy; ++j) cx[((size_t)k * dy + j) * nx + i] = vy[((size_t)i * dz + k) * dy + j];
    axis_interp(cx.data(), (size_t)dz * dy, nx, xc, pts[0].data(), dx, out);             // [dz][dy][dx]
    return GEGS_OK;
}

int gegs_write_interp_dose(const char *root, const float *grid, const uint32_t dimen[3], char *fname, size_t fname_len) {
    if (!root || !grid || !dimen) return fail(GEGS_ERR_ARG, "gegs_write_interp_dose: null argument");
    char name[4096];
    snprintf(name, sizeof name, "%s_Interpolate_%03ux%03ux%03u.bin", root, dimen[0], dimen[1], dimen[2]);
    File f(name, "wb");
    const int32_t d[3] = {(int32_t)dimen[0], (int32_t)dimen[1], (int32_t)dimen[2]};
    const size_t n = (size_t)dimen[0] * dimen[1] * dimen[2];
    if (!f.f || fwrite(d, sizeof d, 1, f.f) != 1 || fwrite(grid, sizeof(float), n, f.f) != n)
        return fail(GEGS_ERR_ARG, "Could not write \"%s\"", name);
    if (fname && fname_len) {
        strncpy(fname, name, fname_len - 1);
        fname[fname_len - 1] = 0;
    }
    return GEGS_OK;
}

int gegs_abs_dose(const float *dose, uint32_t nx, uint32_t ny, uint32_t nz, const float *xc, const float *yc,
                  const float *zc, const float pos[3], int fix_q9, float *value, uint32_t ijk[3], float *voxel_dose) {
    if (!dose || !xc || !yc || !zc || !pos || !value || !ijk || !voxel_dose)
        return fail(GEGS_ERR_ARG, "gegs_abs_dose: null argument");
    if (nx < 3 || ny < 3 || nz < 3) return fail(GEGS_ERR_ARG, "gegs_abs_dose: at least 3 voxels per axis");
    const float px = pos[0], py = pos[1], pz = pos[2];
    uint32_t i = nearest_cell(xc, nx, px), j = nearest_cell(yc, ny, py), k = nearest_cell(zc, nz, pz);
    auto at = [&](uint32_t kk, uint32_t jj, uint32_t ii) { return dose[((size_t)kk * ny + jj) * nx + ii]; };
    auto lin = [](float a, float b, float t, float t0, float t1) {
        float d = (t - t0) * (b - a);
        d = d / (t1 - t0);
        return a + d;
    };
    float a0 = lin(at(k, j, i), at(k + 1, j, i), pz, zc[k], zc[k + 1]);
    float a1 = lin(at(k, j, i + 1), at(k + 1, j, i + 1), pz, zc[k], zc[k + 1]);
    float a2 = lin(at(k, j + 1, i), at(k + 1, j + 1, i), pz, zc[k], zc[k + 1]);
    float a3 = lin(at(k, j + 1, i + 1), at(k + 1, j + 1, i + 1), pz, zc[k], zc[k + 1]);
    a0 = lin(a0, a2, py, yc[j], yc[j + 1]);
    a1 = lin(a1, a3, py, yc[j], yc[j + 1]);
    a0 = lin(a0, a1, px, xc[i], xc[i + 1]);
    if (px > (xc[i] + xc[i + 1]) / 2.0f) ++i;
    if (py > (yc[j] + yc[j + 1]) / 2.0f) ++j;
    // the reference takes the z midpoint from x_bounds[k + 1] (output.c:835, quirk Q9)
    if (!fix_q9 && k + 1 > nx)
        return fail(GEGS_ERR_ARG, "abs_dose: z cell %u needs x bound %u of %u (reference quirk Q9 reads x_bounds for "
                                  "the z midpoint); fix Q9", k, k + 1, nx + 1);
    const float zhi = fix_q9 ? zc[k + 1] : xc[k + 1];
    if (pz > (zc[k] + zhi) / 2.0f) ++k;
    *value = a0;
    ijk[0] = i;
    ijk[1] = j;
    ijk[2] = k;
    *voxel_dose = at(k, j, i);
    return GEGS_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ relative dose (output.c:317-767)
extern "C" {

int gegs_reldose_points(const float *dose, uint32_t nx, uint32_t ny, uint32_t nz, const float *xc, const float *yc,
                        const float *zc, const float start[3], const float sizes[3], const uint32_t dimen[3],
                        float *out) {
    if (!dose || !xc || !yc || !zc || !start || !sizes || !dimen || !out)
        return fail(GEGS_ERR_ARG, "gegs_reldose_points: null argument");
    const float *b[3] = {xc, yc, zc};
    const uint32_t n[3] = {nx, ny, nz};
    for (uint32_t k = 0; k < dimen[2]; ++k)
        for (uint32_t j = 0; j < dimen[1]; ++j)
            for (uint32_t i = 0; i < dimen[0]; ++i) {
                const uint32_t t[3] = {i, j, k};
                long v[3];
                for (int a = 0; a < 3; ++a) {
                    const float p = start[a] + sizes[a] * (float)t[a];
                    const float q = (p - b[a][0]) / (b[a][1] - b[a][0]);
                    v[a] = (long)((double)q + 0.1);               // float + double, truncated toward zero
                    if (v[a] < 0 || v[a] >= (long)n[a])
                        return fail(GEGS_ERR_ARG, "relative dose point (%u, %u, %u) lies outside the phantom on axis %d",
                                    i, j, k, a);
                }
                out[i + (size_t)dimen[0] * j + (size_t)dimen[0] * dimen[1] * k] =
                    dose[((size_t)v[2] * ny + v[1]) * nx + v[0]];
            }
    return GEGS_OK;
}

int gegs_reldose_compare(const float *calc_in, const float *meas_in, uint32_t n, const uint32_t dimen[3],
                         const float start[3], const float sizes[3], const float iso[3], int norm_type, int array_type,
                         float gamma_len, float gamma_dose, float *meas, float *calc, int32_t *pass, uint32_t *total,
                         uint32_t *passed) {
    if (!calc_in || !meas_in || !dimen || !start || !sizes || !iso || !meas || !calc || !pass || !total || !passed)
        return fail(GEGS_ERR_ARG, "gegs_reldose_compare: null argument");
    memcpy(meas, meas_in, sizeof(float) * n);
    memcpy(calc, calc_in, sizeof(float) * n);
    float mmax = 0.0f;
    if (norm_type == 1) {                        // each to its maximum
        float cmax = 0.0f;
        for (uint32_t i = 0; i < n; ++i) {
            if (meas[i] > mmax) mmax = meas[i];
            if (calc[i] > cmax) cmax = calc[i];
        }
        for (uint32_t i = 0; i < n; ++i) {
            meas[i] = meas[i] / mmax;
            calc[i] = calc[i] / cmax;
        }
        mmax = 1.0f;
    } else if (norm_type == 2) {                 // each to its value at the isocentre's point
        for (uint32_t i = 0; i < n; ++i)
            if (meas[i] > mmax) mmax = meas[i];
        long center = 0;
        if (sizes[0] != 0.0f) center += (long)((double)((iso[0] - start[0]) / sizes[0]) + 0.1);
        if (sizes[1] != 0.0f) center += (long)((double)((iso[1] - start[1]) / sizes[1]) + 0.1) * (long)dimen[0];
        if (sizes[2] != 0.0f)
            center += (long)((double)((iso[2] - start[2]) / sizes[2]) + 0.1) * (long)dimen[0] * (long)dimen[1];
        if (center < 0 || center >= (long)n)
            return fail(GEGS_ERR_ARG, "the isocentre's point %ld lies outside the %u relative-dose points", center, n);
        const float mc = meas[center], cc = calc[center];
        for (uint32_t i = 0; i < n; ++i) {
            meas[i] = meas[i] / mc;
            calc[i] = calc[i] / cc;
        }
        mmax = mmax / mc;
    } else {                                     // the measured values x 100
        for (uint32_t i = 0; i < n; ++i) {
            meas[i] = meas[i] * 100.0f;
            if (meas[i] > mmax) mmax = meas[i];
        }
    }
    const float thr = mmax * 0.1f;
    auto near = [&](float c, float m) {
        float v = std::fabs(c - m) / m;
        v = v / gamma_dose;
        v = v * v;
        return std::sqrt((double)v + 0.25 * 0.25 / (double)gamma_len / (double)gamma_len) <= 1.0;
    };
    uint32_t tot = 0, cnt = 0;
    const uint32_t dx = dimen[0];
    for (uint32_t i = 0; i < n; ++i) {
        int p = 0;
        const float m = meas[i];
        if (m > thr) {
            ++tot;
            const bool close = std::fabs(calc[i] - m) / m <= gamma_dose;
            if (array_type == 1) {               // PTW729: the dose difference alone
                if (close) {
                    p = 1;
                    ++cnt;
                } else {
                    p = -1;
                }
            } else {                             // or a gamma-like test against the x and z neighbours
                if (close) p += 1;
                const uint32_t iz = i / dx, ix = i - iz * dx;
                if (ix > 0 && p <= 0 && near(calc[i - 1], m)) p += 1;
                if (ix + 1 < dx && p <= 0 && near(calc[i + 1], m)) p += 1;
                if (iz > 0 && near(calc[i - dx], m)) p += 1;                 // no p <= 0 test (output.c:570)
                if (iz + 1 < dimen[2] && p <= 0 && near(calc[i + dx], m)) p += 1;
                if (p > 0) ++cnt;
                else p = -1;
            }
        }
        pass[i] = p;
    }
    *total = tot;
    *passed = cnt;
    return GEGS_OK;
}

int gegs_write_relative_dose(const char *root, const float *calc, const uint32_t dimen[3], const char *measured_path,
                             const float start[3], const float sizes[3], const float iso[3], int norm_type,
                             int array_type, float gamma_len, float gamma_dose, char *fname, size_t fname_len,
                             char *summary, size_t summary_len) {
    if (!root || !calc || !dimen || !start || !sizes || !iso) return fail(GEGS_ERR_ARG, "gegs_write_relative_dose: null argument");
    char name[4096];
    snprintf(name, sizeof name, "%s_RelativeDose_%03ux%03ux%03u.txt", root, dimen[0], dimen[1], dimen[2]);
    if (fname && fname_len) {
        strncpy(fname, name, fname_len - 1);
        fname[fname_len - 1] = 0;
    }
    const uint32_t n = dimen[0] * dimen[1] * dimen[2];
    File mf(measured_path ? measured_path : "", "r");
    File out(name, "w");
    if (!out.f) return fail(GEGS_ERR_ARG, "Could not write \"%s\"", name);
    auto say = [&](const char *s) {
        if (summary && summary_len) {
            strncpy(summary, s, summary_len - 1);
            summary[summary_len - 1] = 0;
        }
    };
    if (!measured_path || !mf.f) {               // the calculated values alone (output.c:430-436)
        for (uint32_t i = 0; i < n; ++i) fprintf(out.f, "%f\n", (double)calc[i]);
        say("Error: dose data file reading error!");
        return GEGS_OK;
    }
    std::vector<float> meas;
    std::string line;
    while (meas.size() < n && read_line(mf.f, line)) {
        char tok[256];
        if (sscanf(line.c_str(), "%255s", tok) == 1) {
            char *end = nullptr;
            const double v = strtod(tok, &end);
            if (!end || *end) return fail(GEGS_ERR_ARG, "\"%s\": \"%s\" is not a number", measured_path, tok);
            meas.push_back((float)v);
        }
    }
    if (meas.size() < n)
        return fail(GEGS_ERR_ARG, "the measured data file \"%s\" holds %zu of the %u values", measured_path, meas.size(), n);
    std::vector<float> m2(n), c2(n);
    std::vector<int32_t> pass(n);
    uint32_t total = 0, passed = 0;
    const int rc = gegs_reldose_compare(calc, meas.data(), n, dimen, start, sizes, iso, norm_type, array_type, gamma_len,
                                        gamma_dose, m2.data(), c2.data(), pass.data(), &total, &passed);
    if (rc) return rc;
    for (uint32_t i = 0; i < n; ++i) fprintf(out.f, "%f\t %f\t %d\n", (double)m2[i], (double)c2[i], pass[i]);
    const float percent = total ? (float)passed / (float)total : std::nanf("");
    char s[256];
    snprintf(s, sizeof s, "Totally %u points compared, %u passed, percentage = %.0f %%", total, passed,
             (double)(percent * 100.0f));
    fprintf(out.f, "%s\n", s);
    say(s);
    return GEGS_OK;
}

}  // extern "C"
