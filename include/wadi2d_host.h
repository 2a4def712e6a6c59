/* Host-side C-ABI of libwadi2d_host.so: the sub-grid table builder (the inputs of the step) and the seeded noise
 * the synthetic grounds are made of. Plain pointers and sizes; every array is C order. */
#ifndef WADI2D_HOST_H
#define WADI2D_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Table options (config/schema.py:422-442 Tables): rows, local radius on (1, the default) or strips (0), the face
 * depth bounds face_depth_min/max, local_radius_beta and crest_tolerance. */
typedef struct w2d_table_config {
    int32_t rows;
    int32_t local_radius;
    double e_lo, e_hi;
    double beta;
    double crest_tol;
} w2d_table_config;

/* The built tables (tables/builder.py:83-114 GridTables), caller-allocated: zmin, zbar (ny, nx) double; the rest
 * float tile arrays (ntiles, 32, 32 [, rows [, 2]]). */
typedef struct w2d_table_out {
    double *zmin, *zbar;
    float *hc, *dtab;
    float *flx, *frx, *glx, *grx, *efx, *cwx, *ftabx;
    float *fly, *fry, *gly, *gry, *efy, *cwy, *ftaby;
    float *zcx, *zcy;
} w2d_table_out;

/* build_tables (tables/builder.py:131-258) for a uniform grid, no edits and no lining: cell tables
 * (tables/core.py:16-41), face tables (core.py:44-79) with the local-radius conveyance (core.py:82-179), the top
 * window (builder.py:354-371), centre-line beds (builder.py:267-279) and the faces' inverts and natural crests
 * (builder.py:290-308). bed: (bed_ny, bed_nx) pixels of side `pixel` starting at pixel (row_off, col_off) of the
 * padded grid; tiles (ntiles, 2) of (tj, ti), entry 0 the void tile; mask (ny, nx); manning one value
 * (nman_grid 0) or (ny, nx); crest arrays may be null. Returns 0, or 1 on a bad argument. */
int w2d_build_tables(const double* bed, int64_t bed_ny, int64_t bed_nx, int64_t row_off, int64_t col_off,
                     double pixel, int32_t ny, int32_t nx, double dx, const int32_t* tiles, int32_t ntiles,
                     const uint8_t* mask, const double* manning, int32_t nman_grid, const double* crest_xz,
                     const double* crest_xlen, const double* crest_xc, const double* crest_yz,
                     const double* crest_ylen, const double* crest_yc, const w2d_table_config* cfg, int32_t threads,
                     w2d_table_out* out);

/* out[r * nx + c] += amplitude * (value noise, or ridged noise with `ridged`) at the centre of pixel (r, c) of a
 * raster whose north-west corner is (x0, y0): `octaves` octaves from `base_wavelength`, amplitude falling by
 * `persistence` per octave. The reference defines no synthetic ground: this serves the package's seeded
 * workloads. */
void w2d_noise(double* out, int64_t ny, int64_t nx, double x0, double y0, double pixel, uint64_t seed, int32_t octaves,
               double base_wavelength, double persistence, double amplitude, int32_t ridged, int32_t threads);

#ifdef __cplusplus
}
#endif

#endif
