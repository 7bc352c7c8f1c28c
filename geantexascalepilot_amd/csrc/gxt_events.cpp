// gxt_events.cpp -- event front end of libgxt.so (host only): the HEPEVT
// reader and the egamma gun, producing primaries for gxt_submit.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "gxt_data.h"

namespace {

constexpr double kMeMeV = 0.510998910;  // electron mass (MeV), CLHEP value

// Kinetic energy from momentum and mass, the cancellation-free form of
// sqrt(p^2 + m^2) - m used by the reference generators.
double ekin_of(double p, double m) { return p * p / (std::sqrt(p * p + m * m) + m); }

int species_of_pdg(int pdg) {
  if (pdg == 22) return 0;
  if (pdg == 11) return 1;
  if (pdg == -11) return 2;
  return -1;
}

struct HepEntry {
  int ist, id, da1, da2;
  double px, py, pz, m;  // GeV
};

}  // namespace

extern "C" {

int gxt_read_hepevt(const char *path, int max_events, const double vertex[3], uint64_t first, gxt_primary *out,
                    int capacity, int *n_primaries, int *n_events, int *n_unsupported) {
  if (!path || !n_primaries || !n_events || !n_unsupported || (out && capacity < 0))
    return gxt::set_error(GXT_EINVAL, "bad read_hepevt arguments");
  std::ifstream in(path);
  if (!in) return gxt::set_error(GXT_EDATA, std::string("cannot open ") + path);
  const double v[3] = {vertex ? vertex[0] : 0.0, vertex ? vertex[1] : 0.0, vertex ? vertex[2] : 0.0};
  int np = 0, ne = 0, nu = 0;
  std::vector<HepEntry> h;
  std::vector<char> daughter;
  while (max_events < 0 || ne < max_events) {
    int nhep = 0;
    if (!(in >> nhep)) break;  // end of file
    if (nhep < 0) return gxt::set_error(GXT_EDATA, "negative HEPEVT entry count");
    h.resize(nhep);
    for (int i = 0; i < nhep; ++i) {
      HepEntry &e = h[i];
      if (!(in >> e.ist >> e.id >> e.da1 >> e.da2 >> e.px >> e.py >> e.pz >> e.m))
        return gxt::set_error(GXT_EDATA, "truncated HEPEVT event " + std::to_string(ne));
    }
    // daughters (1-based JDAHEP ranges) of the entries that have them
    daughter.assign(nhep, 0);
    for (int i = 0; i < nhep; ++i) {
      if (h[i].da1 <= 0) continue;
      if (h[i].da2 < h[i].da1 || h[i].da2 > nhep)
        return gxt::set_error(GXT_EDATA, "HEPEVT daughter range out of the event in event " + std::to_string(ne));
      for (int j = h[i].da1 - 1; j <= h[i].da2 - 1; ++j)
        if (h[j].ist > 0) daughter[j] = 1;
    }
    for (int i = 0; i < nhep; ++i) {
      const HepEntry &e = h[i];
      // a primary: a daughter with no daughters of its own (the reference's
      // rule), or a stable top-level entry (status 1, nobody's daughter)
      const bool primary = e.da1 == 0 && (daughter[i] || e.ist == 1);
      if (!primary) continue;
      const int sp = species_of_pdg(e.id);
      const double p = std::sqrt(e.px * e.px + e.py * e.py + e.pz * e.pz);
      if (sp < 0 || !(p > 0.0)) {
        ++nu;
        continue;
      }
      if (out) {
        if (np >= capacity) {
          *n_primaries = np;
          return gxt::set_error(GXT_ECAPACITY, "HEPEVT file holds more primaries than capacity");
        }
        gxt_primary &q = out[np];
        std::memset(&q, 0, sizeof(q));
        q.species = sp;
        q.event = ne;
        const double pm = 1.0e3 * p, mm = 1.0e3 * e.m;  // GeV -> MeV
        q.ekin = ekin_of(pm, mm);
        q.dir[0] = e.px / p;
        q.dir[1] = e.py / p;
        q.dir[2] = e.pz / p;
        q.pos[0] = v[0];
        q.pos[1] = v[1];
        q.pos[2] = v[2];
        q.substream = first + (uint64_t)np;
      }
      ++np;
    }
    ++ne;
  }
  *n_primaries = np;
  *n_events = ne;
  *n_unsupported = nu;
  return GXT_OK;
}

int gxt_egamma_gun(uint32_t seed, int n_events, int per_event, double gamma_fraction, double pmin, double pmax,
                   double eta_min, double eta_max, double phi_min, double phi_max, const double vertex[3],
                   uint64_t first, uint64_t n_total, gxt_primary *out) {
  if (!out || n_events < 0 || per_event <= 0 || !(pmax >= pmin) || !(pmin > 0.0))
    return gxt::set_error(GXT_EINVAL, "bad egamma gun arguments");
  const double v[3] = {vertex ? vertex[0] : 0.0, vertex ? vertex[1] : 0.0, vertex ? vertex[2] : 0.0};
  const int n = n_events * per_event;
  for (int i = 0; i < n; ++i) {
    gxt::HostMrg r;
    r.init(seed, n_total + first + (uint64_t)i);
    const bool gamma = r.uniform() < gamma_fraction;
    const double p = pmin + (pmax - pmin) * r.uniform();
    const double eta = eta_min + (eta_max - eta_min) * r.uniform();
    const double phi = phi_min + (phi_max - phi_min) * r.uniform();
    const double theta = 2.0 * std::atan(std::exp(-eta));
    gxt_primary &q = out[i];
    std::memset(&q, 0, sizeof(q));
    q.species = gamma ? 0 : 1;
    q.event = i / per_event;
    q.ekin = ekin_of(p, gamma ? 0.0 : kMeMeV);
    q.dir[0] = std::sin(theta) * std::cos(phi);
    q.dir[1] = std::sin(theta) * std::sin(phi);
    q.dir[2] = std::cos(theta);
    q.pos[0] = v[0];
    q.pos[1] = v[1];
    q.pos[2] = v[2];
    q.substream = first + (uint64_t)i;
  }
  return GXT_OK;
}

}  // extern "C"
