// host_input.cpp -- the run input's head description on the host (declared in
// gegs_host.h): the BEAMnrc component blocks of the .egsinp files (SECJAWS,
// DYNVMLC, PYRAMIDS block and wedge), the .mlcinfo field list with its Varian
// leaf files, and the output-factor table, turned into the gegs_jaws /
// gegs_mlc / gegs_block_wedge descriptions the transport loads.
//
// Host-only C++ (no HIP).  Numbers are read as decimal -> binary64 -> the
// binary32 the geometry holds, and the geometry is assembled in the same
// binary32 operation order as the reference's builders (cited per function),
// so a C++ host and the package's Python readers build identical structures.
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

thread_local std::string g_in_err;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_in_err = buf;
    return code;
}

bool read_lines(const char *path, std::vector<std::string> &lines) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    lines.clear();
    std::string cur;
    int c;
    while ((c = fgetc(f)) != EOF) {
        if (c == '\n') {
            if (!cur.empty() && cur.back() == '\r') cur.pop_back();
            lines.push_back(cur);
            cur.clear();
        } else {
            cur.push_back((char)c);
        }
    }
    if (!cur.empty()) {
        if (cur.back() == '\r') cur.pop_back();
        lines.push_back(cur);
    }
    fclose(f);
    return true;
}

std::string trim(const std::string &s) {
    const char *ws = " \t\r\n\v\f";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return std::string();
    return s.substr(a, s.find_last_not_of(ws) - a + 1);
}

std::vector<std::string> split_ws(const std::string &s) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && isspace((unsigned char)s[i])) ++i;
        size_t j = i;
        while (j < s.size() && !isspace((unsigned char)s[j])) ++j;
        if (j > i) out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

// A whole token as a decimal number (the tokens of these files are plain
// decimals); false when it is not one.
bool to_double(const std::string &t, double &v) {
    if (t.empty()) return false;
    char *end = nullptr;
    errno = 0;
    v = strtod(t.c_str(), &end);
    return end && *end == 0 && end != t.c_str();
}

// DYNVMLC lines: commas count as blanks; the leading numbers of the line
// (a comment may follow them); at least n of them.
bool line_numbers_ws(const std::string &line, size_t n, std::vector<double> &out) {
    std::string s = line;
    for (char &c : s)
        if (c == ',') c = ' ';
    out.clear();
    for (const std::string &t : split_ws(s)) {
        double v;
        if (!to_double(t, v)) break;
        out.push_back(v);
    }
    return out.size() >= n;
}

// PYRAMIDS lines: the first n comma-separated fields, each read as the
// number that starts it (the reference reads ">> value >> ch").
bool line_numbers_csv(const std::string &line, size_t n, std::vector<double> &out) {
    std::string s = line;
    for (char &c : s)
        if (c == '\t') c = ' ';
    out.clear();
    size_t i = 0;
    while (i <= s.size() && out.size() < n) {
        size_t j = s.find(',', i);
        if (j == std::string::npos) j = s.size();
        const std::string t = trim(s.substr(i, j - i));
        if (!t.empty()) {
            double v;
            const std::vector<std::string> w = split_ws(t);
            if (w.empty() || !to_double(w[0], v)) return false;
            out.push_back(v);
        }
        i = j + 1;
    }
    return out.size() == n;
}

void set_name(char dst[GEGS_NAME_LEN], const std::string &s) {
    memset(dst, 0, GEGS_NAME_LEN);
    strncpy(dst, s.c_str(), GEGS_NAME_LEN - 1);
}

struct Leaf {
    double wl, wt, wg, wtip, wts, wbs, ztip, zl, zt, zg, zth, zbh, holepos, zts, zbs;
};

// The 14 cross-section vertices (x, z) of a leaf of one type (read_egsinp.h:37-86).
void leaf_vertices(int kind, const Leaf &lf, float zmin, float zthick, float vx[14], float vz[14]) {
    const double zmax = (double)zmin + (double)zthick;
    double x[14], z[14];
    if (kind == 0) {   // full
        const double a = lf.wt + lf.wl - lf.wg, b = lf.wt + lf.wl, c = lf.wt + lf.wts - lf.wbs, d = lf.wt + lf.wts;
        const double xs[14] = {0, lf.wtip, lf.wtip, a, a, b, b, c, c, d, d, lf.wt, lf.wt, 0};
        const double zs[14] = {lf.ztip, lf.ztip, lf.zl, lf.zl, lf.zg, lf.zg, zmax, zmax, lf.zbs, lf.zbs, lf.zts,
                               lf.zts, lf.zt, lf.zt};
        memcpy(x, xs, sizeof x);
        memcpy(z, zs, sizeof z);
    } else if (kind == 1) {   // target
        const double a = lf.wt + lf.wl - lf.wbs + lf.wts, b = lf.wt + lf.wl - lf.wbs, c = lf.wt + lf.wl,
                     d = lf.wt + lf.wl - lf.wg, e = lf.wt + lf.wl - lf.wg - lf.wtip;
        const double xs[14] = {0, a, a, b, b, c, c, d, d, e, e, lf.wt, lf.wt, 0};
        const double zs[14] = {(double)zmin, (double)zmin, lf.zts, lf.zts, lf.zbs, lf.zbs, lf.zg, lf.zg, lf.ztip,
                               lf.ztip, lf.zl, lf.zl, lf.zt, lf.zt};
        memcpy(x, xs, sizeof x);
        memcpy(z, zs, sizeof z);
    } else {   // isocentre
        const double a = lf.wt + lf.wtip, b = lf.wt + lf.wl - lf.wg, c = lf.wt + lf.wl, d = lf.wts - lf.wbs;
        const double xs[14] = {0, lf.wt, lf.wt, a, a, b, b, c, c, d, d, lf.wts, lf.wts, 0};
        const double zs[14] = {lf.zt, lf.zt, lf.ztip, lf.ztip, lf.zl, lf.zl, lf.zg, lf.zg, zmax, zmax, lf.zbs,
                               lf.zbs, lf.zts, lf.zts};
        memcpy(x, xs, sizeof x);
        memcpy(z, zs, sizeof z);
    }
    for (int i = 0; i < 14; ++i) {
        vx[i] = (float)x[i];
        vz[i] = (float)z[i];
    }
}

}  // namespace

extern "C" {

const char *gegs_input_last_error(void) { return g_in_err.c_str(); }

// ------------------------------------------------------------------ SECJAWS (init.cu:612-689)
int gegs_read_secjaws(const char *path, gegs_jaws *jaws, char medium[GEGS_NAME_LEN]) {
    static const char title[] = "*********** start of CM JAWS with identifier SECJAWS";   // 52 characters
    std::vector<std::string> lines;
    if (!path || !jaws || !medium) return fail(GEGS_ERR_ARG, "gegs_read_secjaws: null argument");
    if (!read_lines(path, lines)) return fail(GEGS_ERR_ARG, "Could not open \"%s\".", path);
    size_t k = 0;
    while (k < lines.size() && lines[k].compare(0, 52, title) != 0) ++k;
    if (k == lines.size()) return GEGS_NOT_FOUND;
    std::vector<double> a, y, x;
    if (k + 7 >= lines.size() || !line_numbers_ws(lines[k + 1], 1, a) || !line_numbers_ws(lines[k + 5], 2, y) ||
        !line_numbers_ws(lines[k + 7], 2, x))
        return fail(2011, "Could not read the SECJAWS component in \"%s\".", path);
    jaws->rmax = (float)a[0];
    jaws->zminy = (float)y[0];
    jaws->zmaxy = (float)y[1];
    jaws->zminx = (float)x[0];
    jaws->zmaxx = (float)x[1];
    set_name(medium, k + 10 < lines.size() ? trim(lines[k + 10]) : std::string());
    return GEGS_OK;
}

// ------------------------------------------------------------------ DYNVMLC (read_egsinp, init.cu:345-611)
int gegs_read_dynvmlc(const char *path, gegs_mlc *m, char medium[GEGS_NAME_LEN]) {
    static const char title[] = "*********** start of CM DYNVMLC";
    std::vector<std::string> lines;
    if (!path || !m || !medium) return fail(GEGS_ERR_ARG, "gegs_read_dynvmlc: null argument");
    if (!read_lines(path, lines)) return fail(GEGS_ERR_ARG, "Could not open \"%s\".", path);
    size_t k = 0;
    while (k < lines.size() && lines[k].compare(0, strlen(title), title) != 0) ++k;
    if (k == lines.size()) return GEGS_NOT_FOUND;
    size_t at = k + 1;
    std::vector<double> v;
    auto next = [&](size_t n, std::vector<double> &out) -> bool {
        if (at >= lines.size()) return false;
        return line_numbers_ws(lines[at++], n, out);
    };
    auto skip = [&](size_t n) { at += n; };
    double rmax, zmin, zthick;
    Leaf full{}, target{}, iso{};
    std::vector<double> fa, fb, fh, fr;
#define GEGS_NEED(n, out) \
    if (!next(n, out)) return fail(2011, "DYNVMLC: expected %d numbers on line %zu of \"%s\"", (int)(n), at, path)
    GEGS_NEED(1, v);
    rmax = v[0];
    skip(2);
    GEGS_NEED(1, v);
    zmin = v[0];
    GEGS_NEED(1, v);
    zthick = v[0];
    GEGS_NEED(6, fa);
    GEGS_NEED(6, fb);
    GEGS_NEED(1, fh);
    GEGS_NEED(2, fr);
    full = Leaf{fa[0], fa[1], fa[2], fa[3], fa[4], fa[5], fb[0], fb[1], fb[2], fb[3], fb[4], fb[5], fh[0], fr[0], fr[1]};
    std::vector<double> ta, tb, th, tc;
    GEGS_NEED(6, ta);
    GEGS_NEED(4, tb);
    GEGS_NEED(1, th);
    GEGS_NEED(4, tc);
    target = Leaf{ta[0], ta[1], ta[2], ta[3], ta[4], ta[5], tc[3], tc[2], tc[0], tc[1], tb[2], tb[3], th[0], tb[0], tb[1]};
    GEGS_NEED(6, fa);
    GEGS_NEED(6, fb);
    GEGS_NEED(1, fh);
    GEGS_NEED(2, fr);
    iso = Leaf{fa[0], fa[1], fa[2], fa[3], fa[4], fa[5], fb[0], fb[1], fb[2], fb[3], fb[4], fb[5], fh[0], fr[0], fr[1]};
    GEGS_NEED(1, v);
    const int n_full = (int)v[0];
    GEGS_NEED(1, v);
    const int n_half = (int)v[0];
    skip(1);
    GEGS_NEED(1, v);
    const double x_start = v[0];
    GEGS_NEED(1, v);
    const double air_gap = v[0];
    skip(1);
    GEGS_NEED(1, v);
    const double rad_end = v[0];
    skip(1);
    struct Opening {
        double neg, pos;
        int num;
    };
    std::vector<Opening> openings;
    int nl = 0;
    while (nl < GEGS_MLC_LEAVES) {
        GEGS_NEED(3, v);
        openings.push_back(Opening{v[0], v[1], (int)v[2]});
        if ((int)v[2] <= 0) return fail(2011, "DYNVMLC: leaf opening count %d in \"%s\"", (int)v[2], path);
        nl += (int)v[2];
    }
#undef GEGS_NEED
    skip(3);
    if (at >= lines.size()) return fail(2011, "DYNVMLC: no medium line in \"%s\"", path);
    set_name(medium, trim(lines[at]));
    if (2 * n_full + n_half != GEGS_MLC_LEAVES || n_half % 2)
        return fail(2011, "DYNVMLC: %d leaves, expected %d", 2 * n_full + n_half, GEGS_MLC_LEAVES);
    if (nl > GEGS_MLC_LEAVES) return fail(2011, "DYNVMLC: openings for %d leaves, expected %d", nl, GEGS_MLC_LEAVES);

    // the builder (init.cu:441-611): focused cross-sections per leaf
    memset(m, 0, sizeof *m);
    const float fzmin = (float)zmin, fzthick = (float)zthick;
    m->zmin = fzmin;
    m->zthick = fzthick;
    m->rmax = (float)rmax;
    float vx[3][14], vz[3][14];
    leaf_vertices(0, full, fzmin, fzthick, vx[0], vz[0]);
    leaf_vertices(1, target, fzmin, fzthick, vx[1], vz[1]);
    leaf_vertices(2, iso, fzmin, fzthick, vx[2], vz[2]);
    const float width[3] = {(float)full.wl, (float)target.wl, (float)iso.wl};
    float X = (float)x_start;
    const float gap = (float)air_gap;
    float x_mid_1 = 0.0f, x_mid_2 = 0.0f;
    for (int i = 0; i < GEGS_MLC_LEAVES; ++i) {
        const int kind = (i < n_full || i >= n_full + n_half) ? 0 : ((i - n_full) % 2 == 0 ? 1 : 2);
        if (i == n_full) x_mid_1 = X;
        if (i == n_full + n_half) x_mid_2 = X;
        for (int j = 0; j < 14; ++j) {
            float q = X + vx[kind][j];
            q = q / fzmin;
            q = q * vz[kind][j];
            m->cross_xz[i * 16 + j].x = q;
            m->cross_xz[i * 16 + j].y = vz[kind][j];
        }
        X = X + width[kind];
        X = X + gap;
    }
    float x_end = X - gap;
    x_end = x_end + (float)full.wg;
    m->x_start = (float)x_start;
    m->x_end = x_end;
    m->x_mid_1 = x_mid_1;
    m->x_mid_2 = x_mid_2;
    m->air_gap = gap;
    m->wl_full = (float)full.wl;
    m->wl_target = (float)target.wl;
    m->wl_isocenter = (float)iso.wl;
    m->wt = (float)full.wt;
    // default openings in the 15th slot of every leaf (init.cu:497-515)
    int li = 0;
    for (const Opening &o : openings)
        for (int c = 0; c < o.num && li < GEGS_MLC_LEAVES; ++c, ++li) {
            m->cross_xz[14 + li * 16].x = (float)o.neg;
            m->cross_xz[14 + li * 16].y = (float)o.pos;
        }
    const float zt = fzmin + fzthick;
    const float xs0 = (float)x_start;
    float b0 = xs0 / fzmin;
    b0 = b0 * zt;
    float b3 = x_end / fzmin;
    b3 = b3 * zt;
    const float bxz[4][2] = {{b0, zt}, {xs0, fzmin}, {x_end, fzmin}, {b3, zt}};
    const float byz[4][2] = {{(float)-rmax, zt}, {(float)-rmax, fzmin}, {(float)rmax, fzmin}, {(float)rmax, zt}};
    for (int q = 0; q < 4; ++q) {
        m->boundary_xz[q].x = bxz[q][0];
        m->boundary_xz[q].y = bxz[q][1];
        m->boundary_yz[q].x = byz[q][0];
        m->boundary_yz[q].y = byz[q][1];
    }
    // the closed leaf pair along the leaf travel: the -rmax edge, bank A's
    // rounded end (tip at y = 0 at mid height), bank B's end, the +rmax edge;
    // 15 points on each end at equal angles, in binary64
    const double dzmin = (double)fzmin, dzthick = (double)fzthick;
    const double zc = dzthick / 2.0;
    const double thmax = std::asin(std::fmin(1.0, zc / rad_end));
    double ang[15];
    const double step = (thmax - -thmax) / 14.0;
    for (int i = 0; i < 15; ++i) ang[i] = (double)i * step + -thmax;
    ang[14] = thmax;
    double ys[15], xs[15];
    for (int i = 0; i < 15; ++i) {
        ys[i] = rad_end * std::sin(ang[i]);
        xs[i] = rad_end * (1.0 - std::cos(ang[i]));
    }
    m->cross_yz[0].x = (float)-rmax;
    m->cross_yz[0].y = (float)(dzmin + dzthick);
    m->cross_yz[1].x = (float)-rmax;
    m->cross_yz[1].y = (float)dzmin;
    for (int i = 0; i < 15; ++i) {
        m->cross_yz[2 + i].x = (float)-xs[i];
        m->cross_yz[2 + i].y = (float)(ys[i] + zc + dzmin);
        m->cross_yz[17 + i].x = (float)xs[14 - i];
        m->cross_yz[17 + i].y = (float)(ys[14 - i] + zc + dzmin);
    }
    m->cross_yz[32].x = (float)rmax;
    m->cross_yz[32].y = (float)dzmin;
    m->cross_yz[33].x = (float)rmax;
    m->cross_yz[33].y = (float)(dzmin + dzthick);
    return GEGS_OK;
}

// ------------------------------------------------------------------ PYRAMIDS block and wedge (init.cu:692-905)
static int find_pyramids(const std::vector<std::string> &lines, const char *ident, size_t &at) {
    std::string title = std::string("*********** start of CM PYRAMIDS with identifier ") + ident;
    title = title.substr(0, 54);
    for (at = 0; at < lines.size(); ++at)
        if (lines[at].compare(0, 54, title) == 0 && lines[at].size() >= title.size()) return 1;
    return 0;
}

int gegs_read_block(const char *path, float coord[10], char medium[GEGS_NAME_LEN]) {
    std::vector<std::string> lines;
    if (!path || !coord || !medium) return fail(GEGS_ERR_ARG, "gegs_read_block: null argument");
    if (!read_lines(path, lines)) return fail(GEGS_ERR_ARG, "Could not open \"%s\".", path);
    size_t i;
    if (!find_pyramids(lines, "BLOCK", i)) return GEGS_NOT_FOUND;
    std::vector<double> v;
    if (i + 7 >= lines.size() || !line_numbers_csv(lines[i + 4], 10, v))
        return fail(2011, "Could not read the block of \"%s\".", path);
    for (int k = 0; k < 10; ++k) coord[k] = (float)v[k];
    set_name(medium, lines[i + 7]);
    return GEGS_OK;
}

int gegs_read_wedge(const char *path, float *rmax, float layers[24], uint32_t *n_layers, char medium[GEGS_NAME_LEN]) {
    std::vector<std::string> lines;
    if (!path || !rmax || !layers || !n_layers || !medium) return fail(GEGS_ERR_ARG, "gegs_read_wedge: null argument");
    if (!read_lines(path, lines)) return fail(GEGS_ERR_ARG, "Could not open \"%s\".", path);
    size_t i;
    if (!find_pyramids(lines, "WEDGE", i)) return GEGS_NOT_FOUND;
    std::vector<double> v;
    if (i + 3 >= lines.size() || !line_numbers_csv(lines[i + 1], 1, v))
        return fail(2011, "Could not read the wedge of \"%s\".", path);
    *rmax = (float)v[0];
    if (!line_numbers_csv(lines[i + 3], 1, v)) return fail(2011, "Could not read the wedge of \"%s\".", path);
    const int n = (int)v[0];
    if (n > 4) return fail(9005, "The number of layers %d is larger than 4!!", n);
    if (n < 1) return fail(9005, "The wedge of \"%s\" has no layer", path);
    for (int j = 0; j < n; ++j) {
        if (i + 4 + j >= lines.size() || !line_numbers_csv(lines[i + 4 + j], 6, v))
            return fail(2011, "Could not read layer %d of the wedge of \"%s\".", j + 1, path);
        for (int q = 0; q < 6; ++q) layers[6 * j + q] = (float)v[q];
    }
    if (i + 4 + n + 2 >= lines.size()) return fail(2011, "The wedge of \"%s\" has no medium line", path);
    *n_layers = (uint32_t)n;
    set_name(medium, lines[i + 4 + n + 2]);
    return GEGS_OK;
}

int gegs_build_block_wedge(const float *block, float wedge_rmax, const float *layers, uint32_t n_layers,
                           gegs_block_wedge *bw) {
    if (!bw) return fail(GEGS_ERR_ARG, "gegs_build_block_wedge: null argument");
    memset(bw, 0, sizeof *bw);
    if (block) {
        // (XFP, zmin), (XBP, zmax), (XBN, zmax), (XFN, zmin) and the same with Y (init.cu:828-844)
        const float *c = block;
        const float xz[4][2] = {{c[2], c[0]}, {c[3], c[1]}, {c[5], c[1]}, {c[4], c[0]}};
        const float yz[4][2] = {{c[6], c[0]}, {c[7], c[1]}, {c[9], c[1]}, {c[8], c[0]}};
        for (int q = 0; q < 4; ++q) {
            bw->block_xz[q] = gegs_float2{xz[q][0], xz[q][1]};
            bw->block_yz[q] = gegs_float2{yz[q][0], yz[q][1]};
        }
        bw->has_block = 1;
    }
    if (n_layers) {
        if (!layers || n_layers > 4) return fail(9005, "The number of layers %u is larger than 4!!", n_layers);
        const float *c = layers;
        const int n = (int)n_layers;
        int start, step, offset;
        // the slanted side of the layer stack (read_egsinp_wedge, init.cu:752-778)
        if (c[2] - c[3] == 0.0f) {
            if (c[4] - c[5] > 0.0f) { start = 4; step = 6; offset = 4; }
            else { start = 6 * n - 1; step = -6; offset = 4; }
        } else if (c[4] - c[5] == 0.0f) {
            if (c[2] - c[3] < 0.0f) { start = 2; step = 6; offset = 2; }
            else { start = 6 * n - 3; step = -6; offset = 2; }
        } else {
            return fail(9005, "neither side of the wedge's first layer is vertical");
        }
        const float first_x = c[start];
        gegs_float2 pts[6];
        int np = 1;
        while (start >= 0 && start < 6 * n) {
            pts[np++] = gegs_float2{c[start], c[start - offset]};
            start += step;
        }
        start = start - step + step / 6;
        pts[np++] = gegs_float2{c[start], c[start - offset]};
        pts[0] = gegs_float2{first_x, c[start - offset]};
        for (int q = 0; q < np; ++q) bw->wedge[q] = pts[q];
        bw->wedge_nver = np;
        const float r = wedge_rmax;
        bw->wedge_bound[0] = gegs_float2{-r, c[0]};
        bw->wedge_bound[1] = gegs_float2{r, c[0]};
        bw->wedge_bound[2] = gegs_float2{r, c[6 * n - 5]};
        bw->wedge_bound[3] = gegs_float2{-r, c[6 * n - 5]};
        bw->has_wedge = 1;
    }
    return GEGS_OK;
}

// ------------------------------------------------------------------ fields (read_mlcinp, init.cu:1377-1502)
int gegs_read_mlcinfo(const char *path, gegs_mlc_field *fields, uint32_t max_fields, uint32_t *n_fields) {
    std::vector<std::string> lines;
    if (!path || !n_fields) return fail(GEGS_ERR_ARG, "gegs_read_mlcinfo: null argument");
    if (!read_lines(path, lines)) return fail(GEGS_ERR_ARG, "Error: MLC info  file reading error! No simulation will be done!");
    if (lines.size() < 2) return fail(2011, "\"%s\" has no field count", path);
    const std::string head = trim(lines[1].substr(0, lines[1].find(',')));
    char *end = nullptr;
    const long n = strtol(head.c_str(), &end, 10);
    if (!end || *end || n < 0) return fail(2011, "\"%s\": bad field count \"%s\"", path, head.c_str());
    if ((size_t)n + 2 > lines.size()) return fail(2011, "\"%s\" lists %ld fields but holds fewer", path, n);
    *n_fields = (uint32_t)n;
    for (long i = 0; i < n && (uint32_t)i < max_fields; ++i) {
        std::vector<std::string> parts;
        const std::string &l = lines[2 + i];
        size_t a = 0;
        while (true) {
            const size_t b = l.find(',', a);
            parts.push_back(trim(l.substr(a, b == std::string::npos ? std::string::npos : b - a)));
            if (b == std::string::npos) break;
            a = b + 1;
        }
        if (parts.size() < 8) return fail(2011, "\"%s\": field %ld has %zu entries", path, i + 1, parts.size());
        double v[6];
        for (int q = 0; q < 6; ++q)
            if (!to_double(parts[1 + q], v[q])) return fail(2011, "\"%s\": field %ld entry %d", path, i + 1, q + 2);
        gegs_mlc_field &f = fields[i];
        f.angle = v[0];
        f.x1 = v[1];
        f.x2 = v[2];
        f.y1 = v[3];
        f.y2 = v[4];
        f.mu = v[5];
        std::string leaf;
        for (size_t q = 7; q < parts.size(); ++q) leaf += (q > 7 ? "," : "") + parts[q];
        leaf = trim(leaf);
        memset(f.leaf_file, 0, sizeof f.leaf_file);
        strncpy(f.leaf_file, leaf.c_str(), sizeof f.leaf_file - 1);
    }
    return GEGS_OK;
}

int gegs_read_leaf_file(const char *path, float positions[2 * GEGS_MLC_LEAVES]) {
    std::vector<std::string> lines;
    if (!path || !positions) return fail(GEGS_ERR_ARG, "gegs_read_leaf_file: null argument");
    if (!read_lines(path, lines)) return fail(GEGS_ERR_ARG, "Could not open the leaf file \"%s\".", path);
    if (lines.size() < 14 + 2 * GEGS_MLC_LEAVES) return fail(2011, "\"%s\" holds fewer than 120 leaves", path);
    for (int j = 0; j < 2 * GEGS_MLC_LEAVES; ++j) {
        const std::string &l = lines[14 + j];
        const std::vector<std::string> w = split_ws(l.size() > 10 ? l.substr(10) : std::string());
        std::string t = w.empty() ? std::string() : w[0];
        size_t a = t.find_first_not_of('='), b = t.find_last_not_of('=');
        t = a == std::string::npos ? std::string() : t.substr(a, b - a + 1);
        double v;
        if (!to_double(t, v)) return fail(2011, "\"%s\": leaf %d", path, j + 1);
        positions[j] = (float)v;
    }
    return GEGS_OK;
}

// ------------------------------------------------------------------ output factors (init.cu:1389-1455)
int gegs_read_output_factors(const char *path, float *sizes, float *factors, uint32_t max, uint32_t *n_out) {
    std::vector<std::string> lines;
    if (!path || !n_out) return fail(GEGS_ERR_ARG, "gegs_read_output_factors: null argument");
    if (!read_lines(path, lines)) return fail(GEGS_ERR_ARG, "Could not open the output-factor file \"%s\".", path);
    if (lines.size() < 2) return fail(2011, "\"%s\" has no entry count", path);
    const std::vector<std::string> h = split_ws(lines[1]);
    char *end = nullptr;
    const long n = h.empty() ? -1 : strtol(h[0].c_str(), &end, 10);
    if (n < 0 || !end || *end || (size_t)n + 2 > lines.size()) return fail(2011, "\"%s\": bad entry count", path);
    *n_out = (uint32_t)n;
    for (long i = 0; i < n && (uint32_t)i < max; ++i) {
        const std::vector<std::string> w = split_ws(lines[2 + i]);
        double a, b;
        if (w.size() < 2 || !to_double(w[0], a) || !to_double(w[1], b))
            return fail(2011, "\"%s\": entry %ld", path, i + 1);
        sizes[i] = (float)a;
        factors[i] = (float)b;
    }
    return GEGS_OK;
}

int gegs_field_output_factor(double x1, double x2, double y1, double y2, const float *sizes, const float *factors,
                             uint32_t n, float *of) {
    if (!of || (n && (!sizes || !factors))) return fail(GEGS_ERR_ARG, "gegs_field_output_factor: null argument");
    // equivalent square of the jaw opening, interpolated in the table, in binary32
    const float dx = (float)x2 - (float)x1, dy = (float)y2 - (float)y1;
    const float a = std::sqrt(dx * dy);
    for (uint32_t j = 0; j + 1 < n; ++j) {
        if (a > sizes[j] && a <= sizes[j + 1]) {
            float t = a - sizes[j];
            t = t * (factors[j + 1] - factors[j]);
            t = t / (sizes[j + 1] - sizes[j]);
            *of = factors[j] + t;
            return GEGS_OK;
        }
    }
    return GEGS_NOT_FOUND;
}

}  // extern "C"
