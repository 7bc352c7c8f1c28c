// gxt_field.h -- host side of the cylindrical (r, z) field map.
#pragma once

#include <string>
#include <vector>

#include "gxt_types.h"

struct gxt_field_map {
  int device = -1;         // HIP device holding dev
  int nr = 0, nz = 0;      // nodes along r and z
  double r0 = 0, z0 = 0;   // first node (mm)
  double dr = 0, dz = 0;   // spacing (mm)
  double bmax = 0;         // largest |B| of the nodes (tesla)
  std::vector<double> v;   // (Br, Bphi, Bz) per node at (ir * nz + iz) * 3
  double *dev = nullptr;   // device copy of v
};

namespace gxt {

// Parses a map file into m (host only; no HIP call).  GXT_EDATA with the
// error message set for a missing, malformed, irregular or truncated file.
int read_field_map(const char *path, gxt_field_map &m);
// Copies m.v to the current HIP device.
int upload_field_map(gxt_field_map &m);
void release_field_map(gxt_field_map &m);
// Device view of a map (v = null for a uniform field b).
DevField dev_field(const gxt_field_map *m, const double b[3], const double origin[3], int stepper);

}  // namespace gxt
