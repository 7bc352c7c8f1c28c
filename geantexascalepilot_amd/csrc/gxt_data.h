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
  std::vector<double> sx;      // merged Sandia intervals of the materials
  std::vector<double> shells;
  std::vector<int> hints;
  bool sb_grid = false;        // params.sbx / sby set from the first element read
};

// Load material `which` (GXT_MAT_*) into params.mat[slot].
int load_material(const std::string &data_dir, int which, double lar_density, HostData &hd, int slot,
                  std::string &err);
void build_geometry(const gxt_config &c, DevGeometry &g);
void build_jump_matrices(uint32_t out[2][64][9]);

// A navigator built by gxt_nav_create: its device view and its touchables
// (host copy; n = their number).
DevNav nav_device_view(const gxt_navigator *nav);
const NavNode *nav_host_nodes(const gxt_navigator *nav, int *n);

// Record msg as the library's last error (gxt_last_error) and return code.
int set_error(int code, const std::string &msg);

// Host MRG32k3a: substream idx of the stream whose six seed words are all
// seed, and its uniform draws (the recurrence, output and jumps of the
// device generator).
struct HostMrg {
  uint64_t s[6];
  void init(uint32_t seed, uint64_t idx);
  double uniform();
};
// Byte-wise jump tables: out[((c * 8 + k) * 256 + d) * kJumpRow + i] = A_c^(d 2^(76+8k)) mod m_c
// for i < 9 (rows padded with zeros to kJumpRow words).
void build_jump_tables(const uint32_t jump[2][64][9], std::vector<uint32_t> &out);

}  // namespace gxt
