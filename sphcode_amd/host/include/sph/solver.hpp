// The step driver: JSON parameters (the reference's keys and defaults,
// src/solver.cpp:155-299), module selection by SPHType, initialisation and
// the integrate() sequence of src/solver.cpp:353-429 over the GPU modules.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "sph/module.hpp"
#include "sph/output.hpp"

namespace sph {

// Parses a JSON document (text) into parameters with the reference's
// defaults; throws SPHException with the reference's messages on missing
// or invalid keys.  Extra keys of this build: "DIM" (1, 2, 3; default 3),
// "precision" ("fp64" or "mixed"), "neighborCapacity".  The required
// "outputDirectory" is returned through output_directory when given.
std::shared_ptr<SPHParameters> read_parameters(const std::string &json_text,
                                               std::string *output_directory = nullptr);
std::shared_ptr<SPHParameters> read_parameter_file(const std::string &path,
                                                   std::string *output_directory = nullptr);

class Solver {
public:
    explicit Solver(std::shared_ptr<SPHParameters> param);
    // Solver::initialize after the initial condition (src/solver.cpp:353-415)
    void initialize(const std::vector<SPHParticle> &particles);
    // In place of initialize: continue from particles saved after a full
    // step (every field of SPHParticle as written) at the saved time, dt and
    // min h/v_sig.  The reference has no restart (its output is lossy text,
    // src/output.cpp:14-64).
    void restart(const std::vector<SPHParticle> &particles, double time, double dt, double h_per_v_sig);
    // Solver::integrate + update_time (src/solver.cpp:417-429)
    void integrate();
    // integrate() until endTime, writing particles and energies through out
    // when given (the loop of src/solver.cpp:300-351); returns the number of
    // steps
    int run(Output *out = nullptr);
    std::shared_ptr<Simulation> simulation() const { return m_sim; }

private:
    std::shared_ptr<SPHParameters> m_param;
    std::shared_ptr<Simulation> m_sim;
    std::shared_ptr<Module> m_timestep, m_pre, m_fforce, m_gforce;
};

}  // namespace sph
