// gxt_data.h -- host-side physics data in the device layout.
#pragma once

#include <string>
#include <vector>

#include "gxt_types.h"

namespace gxt {

struct HostData {
  DevParams params;            // pointers filled in after upload
  std::vector<double> x, y, d2;
  std::vector<double> sb;
  std::vector<double> sandia;
  std::vector<double> shells;
  std::vector<int> hints;
};

// Load material `which` (GXT_MAT_*) into params.mat[slot].
int load_material(const std::string &data_dir, int which, double lar_density, HostData &hd, int slot,
                  std::string &err);
void build_geometry(const gxt_config &c, DevGeometry &g);
void build_jump_matrices(uint32_t out[2][64][9]);
// Byte-wise jump tables: out[((c * 8 + k) * 256 + d) * 9 + i] = A_c^(d 2^(76+8k)) mod m_c.
void build_jump_tables(const uint32_t jump[2][64][9], std::vector<uint32_t> &out);

}  // namespace gxt
