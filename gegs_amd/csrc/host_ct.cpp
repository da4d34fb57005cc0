// host_ct.cpp -- the CT front end on the host (declared in gegs_host.h): a
// DICOM CT series to the egsphant phantom a run reads, as ctcreate builds it
// before the reference's transport starts (ctcreate.c:1342-1600).
//
// Host-only C++ (no HIP).  Cropping and resampling are the library's
// gegs_ct_subsize / gegs_ct_resample; this file reads the series and the ramp
// file, maps CT numbers to media in the reference's binary32 arithmetic and
// writes the egsphant text.  Decimal strings are read as decimal -> binary64
// -> binary32, as the package's Python readers read them, so a C++ host and
// the Python front end build the same phantom bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <unistd.h>

#include "../../include/gegs.h"
#include "../../include/gegs_host.h"

namespace {

thread_local std::string g_ct_err;

constexpr int kRampError = 5007;     // ramps file unreadable or malformed
constexpr int kSeriesError = 5008;   // DICOM series unreadable or malformed

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_ct_err = buf;
    return code;
}

bool read_file(const std::string &path, std::vector<uint8_t> &out) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    out.clear();
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + k);
    const bool ok = !ferror(f);
    fclose(f);
    return ok;
}

// Python's whitespace for ASCII text (str.isspace / str.strip).
bool py_space(char c) {
    return c == ' ' || (c >= '\t' && c <= '\r') || (c >= '\x1c' && c <= '\x1f');
}

std::string strip(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && py_space(s[a])) ++a;
    while (b > a && py_space(s[b - 1])) --b;
    return s.substr(a, b - a);
}

// Lines of a text file split at \n, \r\n and \r (universal newlines), with
// the further separators of str.splitlines when `all_breaks` is set.
std::vector<std::string> text_lines(const std::vector<uint8_t> &d, bool all_breaks) {
    std::vector<std::string> lines;
    std::string cur;
    bool open = false;
    for (size_t i = 0; i < d.size(); ++i) {
        const char c = (char)d[i];
        const bool brk = c == '\n' || c == '\r' ||
                         (all_breaks && (c == '\v' || c == '\f' || c == '\x1c' || c == '\x1d' || c == '\x1e'));
        if (brk) {
            if (c == '\r' && i + 1 < d.size() && d[i + 1] == '\n') ++i;
            lines.push_back(cur);
            cur.clear();
            open = false;
        } else {
            cur.push_back(c);
            open = true;
        }
    }
    if (open) lines.push_back(cur);
    return lines;
}

// A whole token as a decimal (blanks around it allowed); false otherwise.
bool to_double(const std::string &t, double &v) {
    const std::string s = strip(t);
    if (s.empty()) return false;
    char *end = nullptr;
    v = strtod(s.c_str(), &end);
    return end == s.c_str() + s.size();
}

bool to_int(const std::string &t, long long &v) {
    const std::string s = strip(t);
    if (s.empty()) return false;
    char *end = nullptr;
    v = strtoll(s.c_str(), &end, 10);
    return end == s.c_str() + s.size();
}

// ------------------------------------------------------------------ DICOM
constexpr uint32_t TAG_IMAGE_TYPE = 0x00080008u, TAG_IMAGE_POSITION = 0x00200032u, TAG_ROWS = 0x00280010u,
                   TAG_COLUMNS = 0x00280011u, TAG_PIXEL_SPACING = 0x00280030u, TAG_RESCALE_INTERCEPT = 0x00281052u,
                   TAG_RESCALE_SLOPE = 0x00281053u, TAG_PIXEL_DATA = 0x7FE00010u;

uint16_t le16(const uint8_t *p) { return (uint16_t)(p[0] | p[1] << 8); }
uint32_t le32(const uint8_t *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }

bool is_vr(const uint8_t *p, bool &long_len) {
    static const char *vrs[] = {"AE", "AS", "AT", "CS", "DA", "DS", "DT", "FL", "FD", "IS", "LO", "LT", "OB",
                                "OW", "PN", "SH", "SL", "SQ", "SS", "ST", "TM", "UI", "UL", "UN", "US", "UT"};
    for (const char *v : vrs)
        if (p[0] == (uint8_t)v[0] && p[1] == (uint8_t)v[1]) {
            long_len = !strcmp(v, "OB") || !strcmp(v, "OW") || !strcmp(v, "SQ") || !strcmp(v, "UN");
            return true;
        }
    return false;
}

// Every element with a defined, non-zero length in file order, as
// read_elem_vr / read_val walk them (ReadCT_DICOM.c:138-280): a VR is taken
// as explicit when the two bytes after the tag name one, else the element is
// implicit; elements of undefined length are entered, so a sequence's items
// read as further elements; the 128-byte preamble reads as sixteen zero
// elements, after which "DICM" is skipped.  visit(tag, value, length)
// returns false to stop.
template <class Visit>
void for_elements(const std::vector<uint8_t> &d, Visit visit) {
    const size_t n = d.size();
    size_t pos = 0;
    int zeros = 0;
    while (pos + 4 <= n) {
        const uint32_t tag = (uint32_t)le16(&d[pos]) << 16 | le16(&d[pos + 2]);
        pos += 4;
        if (pos + 2 > n) return;
        bool long_len = false;
        uint32_t length;
        if (is_vr(&d[pos], long_len)) {
            if (long_len) {
                if (pos + 8 > n) return;
                length = le32(&d[pos + 4]);
                pos += 8;
            } else {
                if (pos + 4 > n) return;
                length = le16(&d[pos + 2]);
                pos += 4;
            }
        } else {
            if (pos + 4 > n) return;
            length = le32(&d[pos]);
            pos += 4;
        }
        if (length == 0 && tag == 0) ++zeros;
        if (zeros == 16) {
            pos += 4;
            zeros = 0;
        } else if (length > 0 && length < 0xFFFFFFFFu) {
            if (pos + length > n) return;
            if (!visit(tag, &d[pos], length)) return;
            pos += length;
        }
    }
}

std::string value_text(const uint8_t *v, uint32_t len) {
    std::string s((const char *)v, len);
    while (!s.empty() && (s.back() == '\0' || s.back() == ' ')) s.pop_back();
    return s;
}

// string_to_doubles: the first `count` backslash-separated decimals, each finite.
int doubles(const uint8_t *v, uint32_t len, size_t count, double *out) {
    const std::string s = value_text(v, len);
    std::vector<std::string> parts;
    size_t i = 0;
    while (true) {
        const size_t j = s.find('\\', i);
        parts.push_back(s.substr(i, j == std::string::npos ? std::string::npos : j - i));
        if (j == std::string::npos) break;
        i = j + 1;
    }
    if (parts.size() < count) return fail(kSeriesError, "DICOM series: expected %zu values in \"%s\"", count, s.c_str());
    for (size_t k = 0; k < count; ++k)
        if (!to_double(parts[k], out[k]) || !std::isfinite(out[k]))
            return fail(kSeriesError, "DICOM series: \"%s\" is not a finite number", parts[k].c_str());
    return GEGS_OK;
}

int decimal(const uint8_t *v, uint32_t len, float &out) {
    const std::string s = value_text(v, len);
    double x;
    if (!to_double(s, x) || !std::isfinite(x))
        return fail(kSeriesError, "DICOM series: \"%s\" is not a finite number", s.c_str());
    out = (float)x;
    return GEGS_OK;
}

std::string dirname_abs(const std::string &path) {
    std::string p = path;
    if (p.empty() || p[0] != '/') {
        char buf[4096];
        if (!getcwd(buf, sizeof buf)) return ".";
        p = std::string(buf) + "/" + p;
    }
    const size_t k = p.rfind('/');
    return k == 0 ? "/" : p.substr(0, k);
}

std::string join(const std::string &dir, const std::string &name) {
    if (!name.empty() && name[0] == '/') return name;
    if (!dir.empty() && dir.back() == '/') return dir + name;
    return dir + "/" + name;
}

// The binary32 CT number a rescaled value stores to 16 bits: truncated to a
// 32-bit integer (past that range the x86 conversion gives INT_MIN), then the
// low 16 bits, as the reference's float-to-short assignment compiles.
int16_t store_short(float v) {
    const double t = std::trunc((double)v);
    const int32_t i = (std::isfinite(t) && t >= -2147483648.0 && t < 2147483648.0) ? (int32_t)t : INT32_MIN;
    return (int16_t)(uint16_t)((uint32_t)i & 0xFFFFu);
}

struct Slice {
    float z;
    size_t order;
    std::string path;
};

// ------------------------------------------------------------------ ramps
const gegs_ct_ramp kDefaultRamp[4] = {
    {"AIR700ICRU", -974, 0.001f, 0.044f},
    {"LUNG700ICRU", -724, 0.044f, 0.302f},
    {"ICRUTISSUE700ICRU", 101, 0.302f, 1.101f},
    {"ICRPBONE700ICRU", 1976, 1.101f, 2.088f},
};
constexpr int32_t kDefaultLower = -1024;

std::string no_spaces(const std::string &s) {
    std::string o;
    for (char c : s)
        if (c != ' ') o.push_back(c);
    return o;
}

std::vector<std::string> split_commas(const std::string &s) {
    std::vector<std::string> out;
    size_t i = 0;
    while (true) {
        const size_t j = s.find(',', i);
        out.push_back(s.substr(i, j == std::string::npos ? std::string::npos : j - i));
        if (j == std::string::npos) break;
        i = j + 1;
    }
    return out;
}

bool alloc_phantom(gegs_egsphant *ph, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nmed) {
    memset(ph, 0, sizeof *ph);
    ph->nx = nx;
    ph->ny = ny;
    ph->nz = nz;
    ph->nmed = nmed;
    const size_t nv = (size_t)nx * ny * nz;
    ph->names = (char(*)[GEGS_NAME_LEN])calloc(nmed ? nmed : 1, GEGS_NAME_LEN);
    ph->xb = (float *)malloc((nx + 1) * sizeof(float));
    ph->yb = (float *)malloc((ny + 1) * sizeof(float));
    ph->zb = (float *)malloc((nz + 1) * sizeof(float));
    ph->media = (int32_t *)malloc(nv * sizeof(int32_t));
    ph->density = (float *)malloc(nv * sizeof(float));
    return ph->names && ph->xb && ph->yb && ph->zb && ph->media && ph->density;
}

}  // namespace

extern "C" {

const char *gegs_ct_last_error(void) { return g_ct_err.c_str(); }

void gegs_ct_volume_free(gegs_ct_volume *v) {
    if (!v) return;
    free(v->hu);
    v->hu = nullptr;
}

int gegs_dicom_read_series(const char *list_path, gegs_ct_volume *out) {
    if (!list_path || !out) return fail(GEGS_ERR_ARG, "gegs_dicom_read_series: null argument");
    memset(out, 0, sizeof *out);
    std::vector<uint8_t> text;
    if (!read_file(list_path, text)) return fail(kSeriesError, "DICOM series: cannot read %s", list_path);
    const std::string folder = dirname_abs(list_path);
    std::vector<Slice> slices;
    std::vector<float> seen;
    std::vector<uint8_t> data;
    size_t order = 0;
    for (const std::string &name : text_lines(text, false)) {
        if (strip(name).empty()) continue;
        const std::string path = join(folder, name);
        const size_t this_order = order++;
        if (!read_file(path, data)) return fail(kSeriesError, "DICOM series: cannot read %s", path.c_str());
        bool have_z = false;
        float z = 0.0f;
        int rc = GEGS_OK;
        for_elements(data, [&](uint32_t tag, const uint8_t *v, uint32_t len) {
            if (tag == TAG_IMAGE_TYPE) {
                const std::string s((const char *)v, len);
                if (s.find("LOCALIZER") != std::string::npos) return false;
            }
            if (tag == TAG_IMAGE_POSITION) {
                double p[3];
                rc = doubles(v, len, 3, p);
                if (rc == GEGS_OK) {
                    z = (float)(p[2] / 10.0);
                    have_z = true;
                }
                return false;
            }
            return true;
        });
        if (rc) return rc;
        if (!have_z) continue;
        if (std::find(seen.begin(), seen.end(), z) != seen.end()) continue;   // a slice at a z already read
        seen.push_back(z);
        slices.push_back({z, this_order, path});
    }
    if (slices.size() < 2) return fail(kSeriesError, "DICOM series: need at least two CT slices at distinct z");
    std::stable_sort(slices.begin(), slices.end(), [](const Slice &a, const Slice &b) { return a.z < b.z; });
    const float dz = slices[1].z - slices[0].z;
    bool have_rows = false, have_cols = false, have_spacing = false;
    uint32_t rows = 0, cols = 0;
    float offset[3] = {0.0f, 0.0f, 0.0f}, slope = 0.0f, intercept = 0.0f;
    double spacing[2] = {0.0, 0.0};
    std::vector<int16_t> hu;
    for (size_t n = 0; n < slices.size(); ++n) {
        const std::string &path = slices[n].path;
        if (!read_file(path, data)) return fail(kSeriesError, "DICOM series: cannot read %s", path.c_str());
        const uint8_t *pix = nullptr;
        uint32_t pix_len = 0;
        int rc = GEGS_OK;
        for_elements(data, [&](uint32_t tag, const uint8_t *v, uint32_t len) {
            if (n == 0) {
                if (tag == TAG_ROWS || tag == TAG_COLUMNS) {
                    if (len < 2) {
                        rc = fail(kSeriesError, "DICOM series: a rows or columns element shorter than 2 bytes");
                        return false;
                    }
                    (tag == TAG_ROWS ? rows : cols) = le16(v);
                    (tag == TAG_ROWS ? have_rows : have_cols) = true;
                } else if (tag == TAG_IMAGE_POSITION) {
                    double p[3];
                    if ((rc = doubles(v, len, 3, p))) return false;
                    for (int a = 0; a < 3; ++a) offset[a] = (float)(p[a] / 10.0);
                } else if (tag == TAG_PIXEL_SPACING) {
                    if ((rc = doubles(v, len, 2, spacing))) return false;
                    have_spacing = true;
                } else if (tag == TAG_RESCALE_INTERCEPT) {
                    if ((rc = decimal(v, len, intercept))) return false;
                } else if (tag == TAG_RESCALE_SLOPE) {
                    if ((rc = decimal(v, len, slope))) return false;
                }
            }
            if (tag == TAG_PIXEL_DATA) {
                if (len % 2) {
                    rc = fail(kSeriesError, "DICOM series: %s: pixel data of odd length", path.c_str());
                    return false;
                }
                pix = v;
                pix_len = len;
            }
            return true;
        });
        if (rc) return rc;
        if (!have_rows || !have_cols || !have_spacing || !pix)
            return fail(kSeriesError, "DICOM series: %s: missing rows, columns, pixel spacing or pixel data",
                        path.c_str());
        if (rows == 0 || cols == 0) return fail(kSeriesError, "DICOM series: %s: no rows or no columns", path.c_str());
        const size_t plane = (size_t)rows * cols;
        if (pix_len / 2 < plane) return fail(kSeriesError, "DICOM series: %s: short pixel data", path.c_str());
        for (size_t i = 0; i < plane; ++i) hu.push_back((int16_t)le16(pix + 2 * i));
    }
    if (slope != 0.0f && intercept != 0.0f)
        for (int16_t &h : hu) {
            float v = (float)h * slope;
            v = v + intercept;
            h = store_short(v);
        }
    out->size[0] = (int32_t)cols;
    out->size[1] = (int32_t)rows;
    out->size[2] = (int32_t)slices.size();
    out->voxel[0] = (float)(spacing[1] / 10.0);
    out->voxel[1] = (float)(spacing[0] / 10.0);
    out->voxel[2] = dz;
    for (int a = 0; a < 3; ++a) out->offset[a] = (float)((double)offset[a] - (double)out->voxel[a] / 2.0);
    out->hu = (int16_t *)malloc(hu.size() * sizeof(int16_t));
    if (!out->hu) return fail(GEGS_ERR_ARG, "DICOM series: out of host memory");
    memcpy(out->hu, hu.data(), hu.size() * sizeof(int16_t));
    return GEGS_OK;
}

int gegs_read_ct_ramps(const char *path, gegs_ct_ramp ramp[GEGS_CT_MAX_RAMP], uint32_t *n_out, int32_t *ct_lower) {
    if (!path || !ramp || !n_out || !ct_lower) return fail(GEGS_ERR_ARG, "gegs_read_ct_ramps: null argument");
    std::vector<uint8_t> text;
    if (!read_file(path, text)) return fail(kRampError, "CT ramps file %s: cannot be read", path);
    const std::vector<std::string> lines = text_lines(text, true);
    long long n = 0, lower = 0;
    {
        const std::vector<std::string> head = lines.empty() ? std::vector<std::string>() : split_commas(no_spaces(lines[0]));
        if (head.size() < 2 || !to_int(head[0], n) || !to_int(head[1], lower))
            return fail(kRampError, "CT ramps file %s: no \"number of media, lowest CT number\" line", path);
    }
    if (n < 0 || n > GEGS_CT_MAX_RAMP) return fail(5005, "Number of materials out of range (%lld)", n);
    if (n == 0) {
        memcpy(ramp, kDefaultRamp, sizeof kDefaultRamp);
        *n_out = 4;
        *ct_lower = kDefaultLower;
        return GEGS_OK;
    }
    for (long long m = 0; m < n; ++m) {
        const size_t li = 1 + 2 * (size_t)m;
        std::vector<std::string> vals;
        bool ok = li + 1 < lines.size();
        long long up = 0;
        double lo = 0, hi = 0;
        if (ok) {
            for (const std::string &v : split_commas(no_spaces(lines[li + 1])))
                if (!v.empty()) vals.push_back(v);
            ok = vals.size() >= 3 && to_int(vals[0], up) && to_double(vals[1], lo) && to_double(vals[2], hi);
        }
        if (!ok)
            return fail(kRampError,
                        "CT ramps file %s: medium %lld needs a name line and a \"CT upper bound, density low, "
                        "density high\" line", path, m + 1);
        gegs_ct_ramp &r = ramp[m];
        memset(r.name, 0, GEGS_NAME_LEN);
        const std::string nm = strip(lines[li].substr(0, 24));
        strncpy(r.name, nm.c_str(), GEGS_NAME_LEN - 1);
        r.ct_upper = (int32_t)up;
        r.density_lower = (float)lo;
        r.density_upper = (float)hi;
        if (up < lower || r.density_lower < 0.0f || r.density_upper < r.density_lower)
            return fail(5006, "CT no. or density out of range, or density upper bound < density lower bound "
                              "(medium %lld)", m + 1);
    }
    *n_out = (uint32_t)n;
    *ct_lower = (int32_t)lower;
    return GEGS_OK;
}

int gegs_ct_to_medium(const int16_t *hu, size_t n, const gegs_ct_ramp *ramp, uint32_t n_ramp, int32_t ct_lower,
                      int fix_q10, int32_t *media, float *density) {
    if ((n && (!hu || !media || !density)) || !ramp || n_ramp < 1 || n_ramp > GEGS_CT_MAX_RAMP)
        return fail(GEGS_ERR_ARG, "gegs_ct_to_medium: bad argument");
    const gegs_ct_ramp &top = ramp[n_ramp - 1];
    for (size_t v = 0; v < n; ++v) {
        const int32_t ct = hu[v];
        if (ct < ct_lower) {
            media[v] = fix_q10 ? 1 : 0;
            density[v] = fix_q10 ? ramp[0].density_lower : 0.0f;
        } else if (ct > top.ct_upper) {
            media[v] = (int32_t)n_ramp;
            density[v] = top.density_upper;
        } else {
            for (uint32_t i = 0; i < n_ramp; ++i) {
                if (ct > ramp[i].ct_upper) continue;
                const int32_t lo_ct = i == 0 ? ct_lower : ramp[i - 1].ct_upper;
                const float span = ramp[i].density_upper - ramp[i].density_lower;
                float d = (float)(ct - lo_ct) * span;
                d = d / (float)(ramp[i].ct_upper - lo_ct);
                density[v] = ramp[i].density_lower + d;
                media[v] = (int32_t)i + 1;
                break;
            }
        }
    }
    return GEGS_OK;
}

int gegs_ctcreate(const char *list_path, const float thickness[3], const float box_min[3], const float box_max[3],
                  const gegs_ct_ramp *ramp, uint32_t n_ramp, int32_t ct_lower, int fix_q10, gegs_egsphant *out) {
    if (!list_path || !thickness || !box_min || !box_max || !ramp || !out || n_ramp < 1 || n_ramp > GEGS_CT_MAX_RAMP)
        return fail(GEGS_ERR_ARG, "gegs_ctcreate: bad argument");
    memset(out, 0, sizeof *out);
    gegs_ct_volume vol;
    int rc = gegs_dicom_read_series(list_path, &vol);
    if (rc) return rc;
    struct Free {
        gegs_ct_volume *v;
        ~Free() { gegs_ct_volume_free(v); }
    } guard{&vol};
    rc = gegs_ct_subsize(vol.size, vol.hu, vol.voxel, vol.offset, box_min, box_max);
    if (rc) return fail(rc, "gegs_ct_subsize: error %d (CT of %d x %d x %d voxels)", rc, vol.size[0], vol.size[1],
                        vol.size[2]);
    float t[3] = {thickness[0], thickness[1], thickness[2]};
    int32_t o[3];
    rc = gegs_ct_resample_grid(vol.size, vol.voxel, t, o);
    if (rc) return fail(rc, "ResampleCT: error %d", rc);
    if (!alloc_phantom(out, (uint32_t)o[0], (uint32_t)o[1], (uint32_t)o[2], n_ramp)) {
        gegs_egsphant_free(out);
        return fail(GEGS_ERR_ARG, "gegs_ctcreate: out of host memory");
    }
    const size_t nv = (size_t)o[0] * o[1] * o[2];
    std::vector<int16_t> res(nv);
    rc = gegs_ct_resample(vol.size, vol.voxel, vol.hu, vol.offset, t, o, out->xb, out->yb, out->zb, res.data());
    if (!rc) rc = gegs_ct_to_medium(res.data(), nv, ramp, n_ramp, ct_lower, fix_q10, out->media, out->density);
    if (rc) {
        gegs_egsphant_free(out);
        return fail(rc, "gegs_ctcreate: resampling error %d", rc);
    }
    for (uint32_t m = 0; m < n_ramp; ++m) {
        memset(out->names[m], 0, GEGS_NAME_LEN);
        strncpy(out->names[m], ramp[m].name, GEGS_NAME_LEN - 1);
    }
    return GEGS_OK;
}

int gegs_egsphant_write_ctcreate(const char *path, const gegs_egsphant *ph) {
    if (!path || !ph || !ph->names || !ph->xb || !ph->yb || !ph->zb || !ph->media || !ph->density)
        return fail(GEGS_ERR_ARG, "gegs_egsphant_write_ctcreate: null argument");
    for (size_t v = 0, nv = (size_t)ph->nx * ph->ny * ph->nz; v < nv; ++v)
        if (ph->media[v] < 0 || ph->media[v] > 9)
            return fail(GEGS_ERR_ARG, "gegs_egsphant_write_ctcreate: medium %d is not one digit", ph->media[v]);
    FILE *f = fopen(path, "w");
    if (!f) return fail(GEGS_ERR_ARG, "cannot write %s", path);
    const uint32_t nx = ph->nx, ny = ph->ny, nz = ph->nz;
    fprintf(f, "%2d\n", (int)ph->nmed);
    for (uint32_t m = 0; m < ph->nmed; ++m) fprintf(f, "%s\n", ph->names[m]);
    for (uint32_t m = 0; m < ph->nmed; ++m) fprintf(f, " %3.1f", 1.0);
    fprintf(f, "\n %4d %4d %4d\n", (int)nx, (int)ny, (int)nz);
    const float *bounds[3] = {ph->xb, ph->yb, ph->zb};
    const uint32_t counts[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
        for (uint32_t i = 0; i <= counts[a]; ++i) {
            fprintf(f, " %g", (double)bounds[a][i]);
            if (i % 8 == 7) fputc('\n', f);
        }
        if (counts[a] % 8 != 7) fputc('\n', f);
    }
    for (uint32_t k = 0; k < nz; ++k) {
        for (uint32_t j = 0; j < ny; ++j) {
            for (uint32_t i = 0; i < nx; ++i) fprintf(f, "%1d", (int)ph->media[i + nx * (j + (size_t)ny * k)]);
            fputc('\n', f);
        }
        fputc('\n', f);
    }
    for (uint32_t k = 0; k < nz; ++k) {
        for (uint32_t j = 0; j < ny; ++j)
            for (uint32_t i = 0; i < nx; ++i) {
                fprintf(f, "  %g", (double)ph->density[i + nx * (j + (size_t)ny * k)]);
                if ((i + j * nx) % 8 == 7) fputc('\n', f);
            }
        if (((size_t)nx * ny) % 8 != 0) fputc('\n', f);
        fputc('\n', f);
    }
    const bool ok = !ferror(f);
    if (fclose(f) != 0 || !ok) return fail(GEGS_ERR_ARG, "error writing %s", path);
    return GEGS_OK;
}

}  // extern "C"
