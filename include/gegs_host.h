/*
 * gegs_host.h -- host-side C ABI of libgegs_hip.so: the file formats a run
 * reads and writes, restated in C++ so that a C or C++ host drives a run of
 * the reference's input files without Python.  Nothing here touches the GPU;
 * every function runs on a machine without one.
 *
 * Pieces of the reference replaced:
 *   egsphant reader and voxel masses   init.cu:1164-1271 (init_phantom_and_MLC)
 *   region records                     init.cu:909-1051  (init_regions)
 *   egsphsp headers and record reads   read_egsphsp.cpp:60-232
 *   record decoding                    CUDA_EGS.cu:40-106 (load_stack_from_phspfile)
 *   batches of one field               CUDA_EGS.cu:377-411, :462-476, :635-650; init.cu:29-41
 *   field sum and dose file            CUDA_EGS.cu:1712-1735, output.c:873-903 (write_dose)
 * and two files of this library with no counterpart in the reference: the
 * media-table file (the physics tables the HATCH builders produce, so a
 * host loads them without rebuilding them) and the head file (MLC, jaws,
 * block/wedge descriptions and per-field aperture sets).
 *
 * Return values are GEGS_OK or a status code (the reference's 2xxx input and
 * 8xxx phantom codes where it has one); gegs_host_last_error() gives the
 * message of the calling thread's last failure.
 */
#ifndef GEGS_HOST_H
#define GEGS_HOST_H
#include <stddef.h>
#include <stdint.h>
#include "gegs.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GEGS_NAME_LEN 64             /* medium name, NUL-terminated */

/* Reference quirks of the host side with a switch that corrects them.
 * Q4: with one phase-space file the record count comes from <base>.egsphsp1
 *     but the records from <base>_p0.egsphsp1 (read_egsphsp.cpp:113 against
 *     :202-216); the fix reads both from <base>.egsphsp1.
 * Q15: a read running past the end of file i continues at the start of file
 *     i again (read_egsphsp.cpp:208-212); the fix continues in file i + 1.
 * Q16: the last batch of a field holds floor(rest / (nRepeat + 1)) + 1
 *     records, so a field whose histories are a multiple of nRepeat + 1
 *     reads one record past the file and the read refuses the batch
 *     (CUDA_EGS.cu:462-476); the fix transports exactly the field's
 *     histories. */
#define GEGS_HOST_FIX_Q4 1u
#define GEGS_HOST_FIX_Q15 2u
#define GEGS_HOST_FIX_Q16 4u
/* Q9: the absolute-dose voxel's z midpoint is taken with x_bounds[k + 1]
 * (output.c:835); the fix takes z_bounds[k + 1]. */
#define GEGS_HOST_FIX_Q9 8u

const char *gegs_host_last_error(void);

/* ---------------------------------------------------------------- media-table file
 * "GEGSMT01", sizeof(gegs_media_tables), nmed, nmed names of GEGS_NAME_LEN
 * bytes, the table block, an FNV-1a 64 checksum of the block.  names[m] is
 * the medium of table index m. */
int gegs_media_file_write(const char *path, const gegs_media_tables *tables, const char *names);
int gegs_media_file_read(const char *path, gegs_media_tables *tables, char names[GEGS_MXMED][GEGS_NAME_LEN]);

/* ---------------------------------------------------------------- egsphant
 * gegs_egsphant_read restates the reference's reader (init.cu:1164-1251):
 * media count and names, a skipped line, the voxel numbers, nx+1, ny+1, nz+1
 * bounds, one media digit per voxel (x fastest) and one density per voxel,
 * each z slice followed by a blank line; errors 8001-8013 as there.  The
 * arrays are allocated by the library and released by gegs_egsphant_free. */
typedef struct gegs_egsphant {
    uint32_t nx, ny, nz, nmed;
    char (*names)[GEGS_NAME_LEN];    /* nmed medium names */
    float *xb, *yb, *zb;             /* nx + 1, ny + 1, nz + 1 bounds (cm) */
    int32_t *media;                  /* 1-based medium per voxel, 0 = vacuum */
    float *density;                  /* g/cm^3 per voxel */
} gegs_egsphant;

int gegs_egsphant_read(const char *path, gegs_egsphant *out);
void gegs_egsphant_free(gegs_egsphant *ph);

/* Voxel masses (init.cu:1256-1271): |(dx * dy) * dz| * density in binary32,
 * 0 replaced by 1; nx * ny * nz entries. */
int gegs_voxel_mass(const gegs_egsphant *ph, float *mass);

/* Region records (init_regions): region 0 vacuum; regions 1..5 the media
 * other_media[0..4] (air, jaws, MLC, block, wedge) with flags, rhof 1 and the
 * medium's cuts; voxel regions from 6 on with the phantom's media, rhof =
 * density / medium density (1 where the density is 0) and the medium's
 * cuts; vacuum voxels all zero.  Media are found by name in `names` (the
 * media-table file's list); a name not there is error 10029. */
int gegs_region_table(const gegs_egsphant *ph, const gegs_media_tables *tables,
                      const char names[GEGS_MXMED][GEGS_NAME_LEN], const char other_media[5][GEGS_NAME_LEN],
                      uint16_t flags, gegs_region *out /* nx * ny * nz + 6 */);

/* ---------------------------------------------------------------- CT front end
 * The phantom a run builds from a DICOM CT series with "use ctcreate = Yes"
 * (ctcreate.c:1342-1600, called from CUDA_EGS.cu:1327-1383): read the
 * series, crop it to a box, resample it onto transport voxels, map CT numbers
 * to media and densities through a ramp, write the egsphant text the run then
 * reads back.  Errors: status codes with gegs_ct_last_error(). */
#define GEGS_CT_MAX_RAMP 6           /* media in a ramp (MXMED, EGSnrc_Parameters.h:5) */

/* A CT volume: hu[i + nx * (j + ny * k)], x along a row (the column index), y
 * along the columns, z over slices; voxel sizes and the lower corner of voxel
 * (0, 0, 0) in cm.  hu is allocated by the library (gegs_ct_volume_free). */
typedef struct gegs_ct_volume {
    int32_t size[3];
    float voxel[3];
    float offset[3];
    int16_t *hu;
} gegs_ct_volume;

/* readct_dicom (ReadCT_DICOM.c:316-909): the slices named one per line in
 * list_path (relative to the list's folder), little-endian, explicit or
 * implicit VR element by element, with or without the 128-byte preamble;
 * ordered by the z of their image position, localizers and slices at a z
 * already read dropped; slice thickness z[1] - z[0] in binary32; rows,
 * columns, pixel spacing and rescale from the first slice, the rescale
 * applied in binary32 when slope and intercept are both non-zero; lower
 * corner = first pixel centre - half a voxel.  Error 5008 (this library's)
 * for an unreadable file, fewer than two slices, a missing or malformed
 * element or short pixel data; the reference exits there. */
int gegs_dicom_read_series(const char *list_path, gegs_ct_volume *out);
void gegs_ct_volume_free(gegs_ct_volume *v);

/* One medium of a CT ramp: CT numbers up to ct_upper map to it, its density
 * rising linearly from density_lower to density_upper over its CT range. */
typedef struct gegs_ct_ramp {
    char name[GEGS_NAME_LEN];
    int32_t ct_upper;
    float density_lower, density_upper;
} gegs_ct_ramp;

/* The CT ramps file (ctcreate.c:886-1043): "n,lowest CT number", then per
 * medium a name line (its first 24 characters, trimmed) and "CT upper
 * bound, density low, density high, unused"; n = 0 gives the built-in
 * 4-media ramp with lowest CT number -1024 (ctcreate.c:904-929).  Errors
 * 5005 (n outside [0, 6]), 5006 (an upper bound below the lowest CT number,
 * a negative density or densities out of order), 5007 (this library's: the
 * file cannot be read or parsed). */
int gegs_read_ct_ramps(const char *path, gegs_ct_ramp ramp[GEGS_CT_MAX_RAMP], uint32_t *n, int32_t *ct_lower);

/* CTToMedium (ctcreate.c:1052-1104): 1-based medium and density per CT
 * number in binary32; below the ramp vacuum (0, 0.0; quirk Q10,
 * ctcreate.c:1060-1066) or, with fix_q10, the first medium at its lowest
 * density; above the ramp the last medium at its highest density. */
int gegs_ct_to_medium(const int16_t *hu, size_t n, const gegs_ct_ramp *ramp, uint32_t n_ramp, int32_t ct_lower,
                      int fix_q10, int32_t *media, float *density);

/* The whole chain: the series of list_path cropped to [box_min, box_max]
 * (gegs_ct_subsize; all zero = the whole series), resampled on voxels of
 * about `thickness` cm (gegs_ct_resample_grid / gegs_ct_resample), media and
 * densities by gegs_ct_to_medium, the ramp's names as the phantom's media.
 * The phantom is allocated as gegs_egsphant_read allocates it
 * (gegs_egsphant_free). */
int gegs_ctcreate(const char *list_path, const float thickness[3], const float box_min[3], const float box_max[3],
                  const gegs_ct_ramp *ramp, uint32_t n_ramp, int32_t ct_lower, int fix_q10, gegs_egsphant *out);

/* The egsphant text ctcreate writes (ctcreate.c:1249-1311): the media count
 * "%2d", the names, " 1.0" per medium, " %4d" x 3 voxel counts, each axis's
 * bounds " %g" eight to a line, the media digits a row per line with an
 * empty line after each slice, the densities "  %g" eight to a line
 * (counted over a slice) with an empty line after each slice.  The run reads
 * its phantom back from this text, so bounds and densities keep the six
 * significant digits of %g. */
int gegs_egsphant_write_ctcreate(const char *path, const gegs_egsphant *ph);

const char *gegs_ct_last_error(void);

/* ---------------------------------------------------------------- egsphsp
 * The phase-space input of a run: n_files MODE0 files <base>_p<i>.egsphsp1,
 * or one file (host fixes Q4 and Q15 above), read as one record sequence. */
typedef struct gegs_phsp gegs_phsp;

int gegs_phsp_open(const char *base, uint32_t n_files, uint32_t host_fixes, gegs_phsp **out);
uint64_t gegs_phsp_total(const gegs_phsp *ph);
/* Path of file i's header (the file whose record count was read). */
const char *gegs_phsp_header_path(const gegs_phsp *ph, uint32_t i);
/* Records [istart, istart + length) decoded into source particles at the
 * plane z (LATCH bit 30 electron, bit 29 positron, else photon; |E|; W =
 * sqrt(max(0, 1 - U^2 - V^2)) with the sign of WT; |WT|).  Returns
 * GEGS_PHSP_REFUSED where the reference's read returns -1 (istart at or past
 * the total, or fewer than length records left). */
#define GEGS_PHSP_REFUSED 1
int gegs_phsp_read(gegs_phsp *ph, uint64_t istart, uint64_t length, float z, gegs_source *out);
void gegs_phsp_close(gegs_phsp *ph);

/* ---------------------------------------------------------------- batches of a field
 * The reference's batches of a field of `histories` histories over
 * `total_records` records: nRepeat = (histories - 1) / total_records extra
 * uses of every record, batch size from init_stack, n_batch full batches and
 * a last one.  loads[2 i], loads[2 i + 1] = first record and count of every
 * load the read accepts (up to *n_loads; pass max_loads = 0 to learn the
 * count), repeat_counts[q] = records transported with repeat q. */
typedef struct gegs_schedule {
    uint64_t histories;
    uint32_t n_repeat;
    uint32_t batch_size;
    uint64_t n_batch;
    uint64_t last_batch;
    uint64_t n_loads;                /* loads the read accepts */
    uint64_t n_dropped;              /* loads it refuses (Q16) */
    uint64_t dropped_records;
    uint64_t records;                /* records of the accepted loads */
    uint64_t repeat_counts[256];
} gegs_schedule;

int gegs_field_schedule(uint64_t histories, uint64_t total_records, uint32_t host_fixes, gegs_schedule *out,
                        uint64_t *loads, uint64_t max_loads);

/* ---------------------------------------------------------------- dose file
 * write_dose: the plan's deposit (FP64, per voxel) as (float)E / mass *
 * dose_unit in binary32, written to <root>_DoseResults_NNNxNNNxNNN.bin;
 * the file name is returned in fname (at least 32 + strlen(root) bytes). */
int gegs_write_dose(const char *root, uint32_t nx, uint32_t ny, uint32_t nz, const double *edep,
                    const float *mass, float dose_unit, float *dose_out, char *fname, size_t fname_len);

/* ---------------------------------------------------------------- dose outputs of write_dose
 * The coordinates these outputs use are the host's bounds after the upload
 * (init.cu:1289-1294): entry k < n is the centre of voxel k, entry n the last
 * boundary (gegs_host_bounds, n + 1 entries in and out).
 * gegs_interp_dose (interp_dose, output.c:222-312): the dose (binary32,
 * x fastest) linearly interpolated in z, then y, then x onto the points
 * start + sizes * index (dimen per axis); points outside the centres get 0.
 * gegs_write_interp_dose writes <root>_Interpolate_XXXxYYYxZZZ.bin: int32
 * dimensions x, y, z, then the values x fastest.
 * gegs_abs_dose (output_AbsDos, output.c:780-872): the dose at pos
 * (trilinear between voxel centres) and the voxel taken as nearest with its
 * dose; quirk Q9 kept unless fix_q9, and refused where it would read past
 * x_bounds. */
int gegs_host_bounds(const float *bounds, uint32_t n, float *out);
int gegs_interp_dose(const float *dose, uint32_t nx, uint32_t ny, uint32_t nz, const float *xc, const float *yc,
                     const float *zc, const float start[3], const float sizes[3], const uint32_t dimen[3], float *out);
int gegs_write_interp_dose(const char *root, const float *grid, const uint32_t dimen[3], char *fname, size_t fname_len);
int gegs_abs_dose(const float *dose, uint32_t nx, uint32_t ny, uint32_t nz, const float *xc, const float *yc,
                  const float *zc, const float pos[3], int fix_q9, float *value, uint32_t ijk[3], float *voxel_dose);

/* Relative dose (output_RelDos_II, output.c:603-767).  gegs_reldose_points:
 * point (i, j, k) at start + sizes * index takes the dose of voxel
 * int((p - c[0]) / (c[1] - c[0]) + 0.1) per axis (x fastest; a point outside
 * the phantom is refused).  gegs_reldose_compare normalises (norm_type 1:
 * each profile to its maximum, 2: to its value at the isocentre's point,
 * else the measured values x 100) and marks every point whose measured
 * value exceeds 10% of the measured maximum: 1 or more passed, -1 failed,
 * 0 not compared; PTW729 arrays (array_type 1) by the dose difference alone,
 * others also by a gamma-like test against the x and z neighbours 0.25 cm
 * away (gamma_len in cm, gamma_dose a fraction).  gegs_write_relative_dose
 * writes <root>_RelativeDose_XXXxYYYxZZZ.txt ("measured\t calculated\t pass"
 * rows and the pass-rate line, or the calculated values alone when the
 * measured file cannot be opened) and returns the line to log in summary. */
int gegs_reldose_points(const float *dose, uint32_t nx, uint32_t ny, uint32_t nz, const float *xc, const float *yc,
                        const float *zc, const float start[3], const float sizes[3], const uint32_t dimen[3],
                        float *out);
int gegs_reldose_compare(const float *calc, const float *measured, uint32_t n, const uint32_t dimen[3],
                         const float start[3], const float sizes[3], const float isocenter[3], int norm_type,
                         int array_type, float gamma_len, float gamma_dose, float *measured_out, float *calc_out,
                         int32_t *pass, uint32_t *total, uint32_t *passed);
int gegs_write_relative_dose(const char *root, const float *calc, const uint32_t dimen[3], const char *measured_path,
                             const float start[3], const float sizes[3], const float isocenter[3], int norm_type,
                             int array_type, float gamma_len, float gamma_dose, char *fname, size_t fname_len,
                             char *summary, size_t summary_len);

/* ---------------------------------------------------------------- head file
 * "GEGSHD01": which of MLC, jaws and block/wedge are present, their
 * descriptions and n_sets aperture sets (120 leaf positions and x1 x2 y1 y2
 * jaw openings per set). */
typedef struct gegs_head {
    uint32_t has_mlc, has_jaws, has_block_wedge, n_sets;
    gegs_mlc mlc;
    gegs_jaws jaws;
    gegs_block_wedge block_wedge;
    float *leaf_positions;           /* n_sets x 120 (has_mlc) */
    float *jaw_positions;            /* n_sets x 4 (has_jaws) */
} gegs_head;

int gegs_head_file_write(const char *path, const gegs_head *head);
/* Allocates leaf_positions / jaw_positions; release with gegs_head_free. */
int gegs_head_file_read(const char *path, gegs_head *head);
void gegs_head_free(gegs_head *head);

/* ---------------------------------------------------------------- head inputs of a run
 * The BEAMnrc component blocks of an .egsinp file, the .mlcinfo field list,
 * the Varian leaf files and the output-factor table.  Numbers are read as
 * decimal -> binary64 -> binary32; the geometry is built in the reference's
 * binary32 operation order.  A file without the component asked for returns
 * GEGS_NOT_FOUND (the reference then keeps that module's defaults); errors
 * give a status code and gegs_input_last_error(). */
#define GEGS_NOT_FOUND 1

const char *gegs_input_last_error(void);

/* JAWS component SECJAWS (read_egsinp_secjaws, init.cu:612-689): RMAX, the Y
 * and X pairs' z ranges; medium = the jaw medium's name. */
int gegs_read_secjaws(const char *egsinp, gegs_jaws *jaws, char medium[GEGS_NAME_LEN]);

/* DYNVMLC component (read_egsinp, init.cu:345-611): the three leaf types, the
 * leaf order, the default openings and the rounded leaf ends, built into the
 * focused cross-sections of gegs_mlc; medium = the leaf medium. */
int gegs_read_dynvmlc(const char *egsinp, gegs_mlc *mlc, char medium[GEGS_NAME_LEN]);

/* PYRAMIDS components BLOCK (10 values: zmin, zmax, XFP, XBP, XFN, XBN, YFP,
 * YBP, YFN, YBN; init.cu:793-905) and WEDGE (RMAX and 1-4 layers of zmin,
 * zmax, XFP, XBP, XFN, XBN; init.cu:692-791), and the gegs_block_wedge built
 * from them (block or layers may be absent: null / n_layers 0). */
int gegs_read_block(const char *egsinp, float coord[10], char medium[GEGS_NAME_LEN]);
int gegs_read_wedge(const char *egsinp, float *rmax, float layers[24], uint32_t *n_layers,
                    char medium[GEGS_NAME_LEN]);
int gegs_build_block_wedge(const float *block_coord, float wedge_rmax, const float *layers, uint32_t n_layers,
                           gegs_block_wedge *out);

/* One field of a .mlcinfo list (read_mlcinp, init.cu:1377-1502): gantry
 * angle (degrees), jaw opening x1 x2 y1 y2 (cm at the isocentre), monitor
 * units and the leaf file as written. */
typedef struct gegs_mlc_field {
    double angle, x1, x2, y1, y2, mu;
    char leaf_file[512];
} gegs_mlc_field;

/* Fields of the list (*n_fields = the count it declares; at most max_fields
 * are written). */
int gegs_read_mlcinfo(const char *path, gegs_mlc_field *fields, uint32_t max_fields, uint32_t *n_fields);
/* A Varian leaf file: 14 header lines, then 120 positions read after the
 * first 10 characters of their lines (bank A leaves 1-60, then bank B). */
int gegs_read_leaf_file(const char *path, float positions[120]);

/* Output-factor table (init.cu:1389-1412): a header line, the entry count,
 * then "field size  output factor" per line (*n = the count; at most max
 * entries written).  gegs_field_output_factor interpolates it at the
 * equivalent square sqrt((x2 - x1)(y2 - y1)) of a jaw opening, in binary32
 * (init.cu:1430-1455); GEGS_NOT_FOUND outside the table. */
int gegs_read_output_factors(const char *path, float *sizes, float *factors, uint32_t max, uint32_t *n);
int gegs_field_output_factor(double x1, double x2, double y1, double y2, const float *sizes, const float *factors,
                             uint32_t n, float *output_factor);

#ifdef __cplusplus
}
#endif
#endif
