// Full-precision particle snapshot.  The reference writes lossy text
// snapshots (src/output.cpp:14-64); this binary form keeps every FP64 bit
// so that runs can be compared and restarted.
//
// Layout, little endian:
//   char    magic[8]  = "SPHSNAP2"
//   int32   dim, int32 reserved (0)
//   int64   n
//   double  time, dt, h_per_v_sig (the last is absent in "SPHSNAP1" files)
//   then the fields of SPHParticle in this order, each as one array over
//   the particles: pos, vel, vel_p, acc (n*dim doubles, [i*dim+k]); mass,
//   dens, pres, ene, ene_p, dene, sml, sound, balsara, alpha, gradh, phi
//   (n doubles); id, neighbor (n int32).
#pragma once

#include <string>
#include <vector>

#include "sph/simulation.hpp"

namespace sph {

struct Snapshot {
    int dim = 3;
    double time = 0.0, dt = 0.0, h_per_v_sig = 0.0;
    std::vector<SPHParticle> particles;
};

void write_snapshot(const std::string &path, const Snapshot &snap);
Snapshot read_snapshot(const std::string &path);  // throws SPHException

}  // namespace sph
