/*
 * phaser_ply.h -- PLY point-cloud ingestion for the registration path
 * (libphaser_host.so).
 *
 * Replaces data::PlyHelper::readPlyFromFile (ply-helper.cc:25-65, over the
 * tinyply parser) and model::PointCloud::parsePlyPointCloud
 * (point-cloud.cc:313-349): the "vertex" element's x, y, z and the
 * intensity, reflectivity and ambient fields, named by the flags
 * phaser_ply_intensity_str / _reflectivity_str / _ambient_str /
 * _range_str (ply-helper.cc:8-22).  ASCII, binary little-endian and
 * binary big-endian files; every PLY scalar type (values are converted to
 * double, as tinyply does for the reference's std::vector<double>, then to
 * the cloud's float); list properties and other elements are skipped.
 *
 * As in the reference, x, y, z and the intensity field are required (its
 * CHECK_GT aborts; here -1), reflectivity, ambient and range are optional
 * (missing: a zero channel, the reference's empty vector).  The range field
 * is read but the registration path recomputes the range from the
 * coordinates (spherical-projection.cc:22-31), as the reference does.
 */
#ifndef PHASER_PLY_H_
#define PHASER_PLY_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field names; NULL members take the reference's defaults. */
typedef struct {
  const char* intensity;    /* phaser_ply_intensity_str, "intensity" */
  const char* reflectivity; /* phaser_ply_reflectivity_str, "reflectivity" */
  const char* ambient;      /* phaser_ply_ambient_str, "noise" */
  const char* range;        /* phaser_ply_range_str, "range" */
} phz_ply_fields;

typedef struct {
  int64_t n;        /* points: vertices, x y z intensity each */
  float* xyzi;      /* n x 4: x, y, z, intensity (the PointXYZI cloud) */
  float* extra;     /* n x 2: reflectivity, ambient (0 where the field is missing) */
  double* range;    /* n values of the range field, or NULL when missing */
  int32_t has_reflectivity, has_ambient, has_range;
  int32_t reserved;
} phz_ply_cloud;

void phz_ply_default_fields(phz_ply_fields* fields);
/* Reads path into out (free with phz_ply_free).  Returns 0, or -1 with
 * phz_ply_last_error() set (unreadable file, malformed header or body,
 * missing x/y/z or intensity). */
int phz_ply_read(const char* path, const phz_ply_fields* fields, phz_ply_cloud* out);
void phz_ply_free(phz_ply_cloud* cloud);
const char* phz_ply_last_error(void);
/* Writes n points (x, y, z, intensity) as an ASCII PLY, the form of
 * pcl::io::savePLYFileASCII of a PointXYZI cloud that the reference's
 * driver writes the registered cloud with (driver.cc:17-23); floats with 9
 * significant digits, which read back exactly.  Returns 0 or -1. */
int phz_ply_write_xyzi(const char* path, int64_t n, const float* xyzi);

#ifdef __cplusplus
}
#endif
#endif /* PHASER_PLY_H_ */
