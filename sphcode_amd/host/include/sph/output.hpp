// Run output in the manner of the reference's Output (src/output.cpp:14-90):
// energy.dat in the output directory, one line of time, kinetic, thermal,
// potential and total energy per call, plus the total momentum (DIM
// columns; the reference writes none); particle output as full-precision
// snapshots (snapshot_NNNN.snap, the format of sph/snapshot.hpp) instead
// of the reference's lossy text.
#pragma once

#include <fstream>
#include <memory>
#include <string>

#include "sph/simulation.hpp"

namespace sph {

class Output {
public:
    // Creates the directory if needed and opens energy.dat with its header
    // line (src/output.cpp:34-41).
    explicit Output(const std::string &directory);
    // Output::output_particle: the state as snapshot_<count>.snap.
    void output_particle(Simulation &sim);
    // Output::output_energy: the FP64 device sums (src/output.cpp:66-90).
    void output_energy(const Simulation &sim);

private:
    std::string m_dir;
    std::ofstream m_energy;
    int m_count = 0;
};

}  // namespace sph
