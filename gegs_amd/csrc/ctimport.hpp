// ctimport.hpp -- host side of the CT front end: crop a CT volume to a box
// and resample it onto the transport voxels, in the binary32 arithmetic and
// summation order of ctcreate (SubsizeCT, ctcreate.c:40-274; ResampleCT,
// ctcreate.c:277-823), so that the resampled CT numbers are the reference's
// integers.  Volumes are x-fastest: v[i + nx * (j + ny * k)].
#pragma once
#include <cstdint>
#include <vector>

namespace gegs {
namespace ct {

// voxel boundaries accumulated one voxel at a time in binary32
inline std::vector<float> bounds(float lo, float width, int n) {
    std::vector<float> b((size_t)n + 1);
    b[0] = lo;
    for (int i = 1; i <= n; ++i) b[i] = b[i - 1] + width;
    return b;
}

// SubsizeCT's choice of voxel range along one axis for the requested
// [lo, hi]: an empty or all-outside request keeps the whole axis, a partial
// one is clipped to the data; the range runs from the voxel holding lo to
// the last voxel whose closed extent holds hi.
inline void sub_range(const std::vector<float> &b, int n, float lo, float &hi, bool whole, int &first, int &last) {
    if (whole) {
        lo = b[0];
        hi = b[n];
    }
    if ((lo < b[0] && hi < b[0]) || (hi > b[n] && lo > b[n])) {
        lo = b[0];
        hi = b[n];
    }
    if (hi <= lo) {
        lo = b[0];
        hi = b[n];
    }
    if (lo < b[0]) lo = b[0];
    if (hi > b[n]) hi = b[n];
    first = 0;
    last = n - 1;
    for (int i = 0; i < n; ++i) {
        if (b[i] <= lo && b[i + 1] > lo) first = i;
        if (b[i] <= hi && b[i + 1] >= hi) last = i;
    }
}

// Returns the new size and offset; hu is compacted in place.
inline void subsize(int size[3], int16_t *hu, const float voxel[3], float offset[3], const float sub_min[3],
                    const float sub_max[3]) {
    const bool whole = sub_min[0] == 0.0f && sub_max[0] == 0.0f && sub_min[1] == 0.0f && sub_max[1] == 0.0f &&
                       sub_min[2] == 0.0f && sub_max[2] == 0.0f;
    int first[3], last[3];
    for (int a = 0; a < 3; ++a) {
        const std::vector<float> b = bounds(offset[a], voxel[a], size[a]);
        float hi = sub_max[a];
        sub_range(b, size[a], sub_min[a], hi, whole, first[a], last[a]);
        // the reference's special case of a box ending on the lower face of
        // the last z voxel keeps that voxel
        if (a == 2 && size[2] >= 1 && b[size[2] - 1] == hi) last[2] = size[2] - 1;
    }
    const int nx = size[0], ny = size[1];
    size_t o = 0;
    for (int k = first[2]; k <= last[2]; ++k)
        for (int j = first[1]; j <= last[1]; ++j)
            for (int i = first[0]; i <= last[0]; ++i) hu[o++] = hu[i + (size_t)nx * (j + (size_t)ny * k)];
    for (int a = 0; a < 3; ++a) {
        offset[a] = offset[a] + voxel[a] * (float)first[a];
        size[a] = last[a] - first[a] + 1;
    }
}

// ResampleCT's output grid: whole voxels of about `thickness` spanning the
// CT; returns 0 or the reference's input error (5002 non-positive, 5003
// larger than the CT, 5004 more than max_cells voxels along an axis).
inline int resample_grid(const int size[3], const float voxel[3], float thickness[3], int out[3], int max_cells) {
    for (int a = 0; a < 3; ++a)
        if (thickness[a] <= 0.0f) return 5002;
    for (int a = 0; a < 3; ++a)
        if (thickness[a] > (float)size[a] * voxel[a]) return 5003;
    for (int a = 0; a < 3; ++a)
        if (thickness[a] < (float)size[a] * voxel[a] / (float)max_cells) return 5004;
    for (int a = 0; a < 3; ++a) {
        out[a] = (int)((float)size[a] * voxel[a] / thickness[a]);
        thickness[a] = (float)size[a] * voxel[a] / (float)out[a];
    }
    return 0;
}

// Fractions of the output voxels [first, last] along one axis covered by CT
// voxel c (its extent measured from the first output bound, as the
// reference does).  first and last keep their previous values when no
// output voxel matches (the top CT voxel's upper face can lie above the
// accumulated last bound by a rounding), as the reference's do.
inline void axis_weights(const std::vector<float> &xb, int nout, float ctw, int c, int &first, int &last,
                         std::vector<float> &w) {
    const float lo = xb[0] + ctw * (float)c;
    const float hi = xb[0] + ctw * (float)(c + 1);
    for (int i = 0; i < nout; ++i) {
        if (xb[i] <= lo && xb[i + 1] >= lo) first = i;
        if (xb[i] <= hi && xb[i + 1] >= hi) {
            last = i;
            break;
        }
    }
    if (first == last) {
        w[first] = ctw / (xb[first + 1] - xb[first]);
        return;
    }
    for (int i = first; i <= last; ++i) {
        if (xb[i] >= lo && xb[i + 1] <= hi) w[i] = 1.0f;
        else if (xb[i] <= lo && xb[i + 1] <= hi) w[i] = (xb[i + 1] - lo) / (xb[i + 1] - xb[i]);
        else w[i] = (hi - xb[i]) / (xb[i + 1] - xb[i]);
    }
}

// Volume-weighted CT numbers on the output grid, accumulated in binary32 in
// the reference's order (CT voxel i, then j, then k; output voxels i, j, k
// inside), converted to integers by truncation.
inline void resample(const int size[3], const float voxel[3], const int16_t *hu, const float offset[3],
                     const float thickness[3], const int out[3], float *xb_out, float *yb_out, float *zb_out,
                     int16_t *res) {
    const std::vector<float> xb = bounds(offset[0], thickness[0], out[0]);
    const std::vector<float> yb = bounds(offset[1], thickness[1], out[1]);
    const std::vector<float> zb = bounds(offset[2], thickness[2], out[2]);
    for (int i = 0; i <= out[0]; ++i) xb_out[i] = xb[i];
    for (int i = 0; i <= out[1]; ++i) yb_out[i] = yb[i];
    for (int i = 0; i <= out[2]; ++i) zb_out[i] = zb[i];
    const size_t nout = (size_t)out[0] * out[1] * out[2];
    std::vector<float> acc(nout, 0.0f);
    std::vector<float> wx(out[0], 0.0f), wy(out[1], 0.0f), wz(out[2], 0.0f);
    int i0 = 0, i1 = 0, j0 = 0, j1 = 0, k0 = 0, k1 = 0;
    for (int ic = 0; ic < size[0]; ++ic) {
        axis_weights(xb, out[0], voxel[0], ic, i0, i1, wx);
        for (int jc = 0; jc < size[1]; ++jc) {
            axis_weights(yb, out[1], voxel[1], jc, j0, j1, wy);
            for (int kc = 0; kc < size[2]; ++kc) {
                axis_weights(zb, out[2], voxel[2], kc, k0, k1, wz);
                const float v = (float)hu[ic + (size_t)size[0] * (jc + (size_t)size[1] * kc)];
                for (int i = i0; i <= i1; ++i)
                    for (int j = j0; j <= j1; ++j)
                        for (int k = k0; k <= k1; ++k)
                            acc[i + (size_t)out[0] * (j + (size_t)out[1] * k)] += v * wx[i] * wy[j] * wz[k];
            }
        }
    }
    for (size_t v = 0; v < nout; ++v) res[v] = (int16_t)(int)acc[v];
}

}  // namespace ct
}  // namespace gegs
