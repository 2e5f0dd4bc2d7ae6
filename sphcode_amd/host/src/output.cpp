#include "sph/output.hpp"

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>

#include "sph/exception.hpp"
#include "sph/snapshot.hpp"

namespace sph {

Output::Output(const std::string &directory) : m_dir(directory) {
    if (::mkdir(m_dir.c_str(), 0755) != 0 && errno != EEXIST)
        throw SPHException("cannot create output directory " + m_dir);
    const std::string name = m_dir + "/energy.dat";
    m_energy.open(name);
    if (!m_energy) throw SPHException("cannot open " + name);
    m_energy << "# time kinetic thermal potential total momentum\n";
}

void Output::output_particle(Simulation &sim) {
    char name[32];
    std::snprintf(name, sizeof(name), "/snapshot_%04d.snap", m_count);
    Snapshot s;
    s.dim = sim.params().dim;
    s.time = sim.get_time();
    s.dt = sim.get_dt();
    s.h_per_v_sig = sim.get_h_per_v_sig();
    s.particles = sim.particles_view();
    write_snapshot(m_dir + name, s);
    ++m_count;
}

void Output::output_energy(const Simulation &sim) {
    const auto e = sim.energy();
    char line[512];
    int len = std::snprintf(line, sizeof(line), "%.17g %.17g %.17g %.17g %.17g", sim.get_time(), e[0], e[1], e[2],
                            e[3]);
    for (int k = 0; k < sim.params().dim; ++k)
        len += std::snprintf(line + len, sizeof(line) - len, " %.17g", e[4 + k]);
    m_energy << line << std::endl;
}

}  // namespace sph
