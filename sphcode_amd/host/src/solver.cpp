// Solver: module selection, initialisation and the step sequence
// (src/solver.cpp:300-429), over the GPU modules.

#include <cmath>

#include "sph/exception.hpp"
#include "sph/solver.hpp"

namespace sph {

Solver::Solver(std::shared_ptr<SPHParameters> param) : m_param(std::move(param)) {
    if (!m_param) throw SPHException("Solver: no parameters");
    // module selection by SPHType (src/solver.cpp:357-370)
    m_timestep = std::make_shared<gpu::TimeStep>();
    if (m_param->type == SPHType::SSPH) {
        m_pre = std::make_shared<gpu::PreInteraction>();
        m_fforce = std::make_shared<gpu::FluidForce>();
    } else if (m_param->type == SPHType::DISPH) {
        m_pre = std::make_shared<gpu::disph::PreInteraction>();
        m_fforce = std::make_shared<gpu::disph::FluidForce>();
    } else {
        m_pre = std::make_shared<gpu::gsph::PreInteraction>();
        m_fforce = std::make_shared<gpu::gsph::FluidForce>();
    }
    m_gforce = std::make_shared<gpu::GravityForce>();
    m_timestep->initialize(m_param);
    m_pre->initialize(m_param);
    m_fforce->initialize(m_param);
    m_gforce->initialize(m_param);
}

// src/solver.cpp:392-414: alpha, Balsara and sound speed of the initial
// condition, first tree, then the three force modules once.
void Solver::initialize(const std::vector<SPHParticle> &particles) {
    m_sim = std::make_shared<Simulation>(m_param);
    std::vector<SPHParticle> p = particles;
    const double gamma = m_param->physics.gamma;
    const double c_sound = gamma * (gamma - 1.0);
    for (auto &pi : p) {
        pi.alpha = m_param->av.alpha;
        pi.balsara = 1.0;
        pi.sound = std::sqrt(c_sound * pi.ene);
    }
    m_sim->set_particles(p);
    m_sim->make_tree();
    m_pre->calculation(m_sim);
    m_fforce->calculation(m_sim);
    m_gforce->calculation(m_sim);
}

void Solver::restart(const std::vector<SPHParticle> &particles, double time, double dt, double h_per_v_sig) {
    m_sim = std::make_shared<Simulation>(m_param);
    m_sim->set_particles(particles);
    m_sim->restart(time, dt, h_per_v_sig);
}

// src/solver.cpp:417-429 and the time update of Solver::run (:327)
void Solver::integrate() {
    if (!m_sim) throw SPHException("Solver::integrate before initialize");
    void *ctx = m_sim->device();
    m_timestep->calculation(m_sim);
    check_status(sphgpu_predict(ctx), "sphgpu_predict");
    m_sim->make_tree();
    m_pre->calculation(m_sim);
    m_fforce->calculation(m_sim);
    m_gforce->calculation(m_sim);
    check_status(sphgpu_correct(ctx), "sphgpu_correct");
    m_sim->update_time();
}

// the loop of Solver::run (src/solver.cpp:300-351): integrate until
// endTime; with an Output, the particles and the energies at the start and
// after every step that passes the next outputTime / energyTime mark
int Solver::run(Output *out) {
    if (!m_sim) throw SPHException("Solver::run before initialize");
    double t_out = m_param->time.output;
    double t_ene = m_param->time.energy;
    if (out) {
        out->output_particle(*m_sim);
        out->output_energy(*m_sim);
    }
    int loop = 0;
    double t = m_sim->get_time();
    while (t < m_param->time.end) {
        integrate();
        ++loop;
        t = m_sim->get_time();
        if (out && t > t_out) {
            out->output_particle(*m_sim);
            t_out += m_param->time.output;
        }
        if (out && t > t_ene) {
            out->output_energy(*m_sim);
            t_ene += m_param->time.energy;
        }
    }
    return loop;
}

}  // namespace sph
