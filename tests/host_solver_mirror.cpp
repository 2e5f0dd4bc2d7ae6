// A solver in the shape of the reference's Solver (src/solver.cpp:353-474),
// written against this build's host API: module selection by SPHType
// (:357-370), the GSPH vector arrays (:373-385), the modules' initialize
// (:387-390), the initial alpha / Balsara / sound loop over the host
// particles (:392-404), the tree's resize and make (:406-410), the three
// force modules once (:412-414); per step TimeStep, predict as a host loop
// over the particles (:431-456), make_tree, the force modules, and correct
// as a host loop (:458-474).  Test code: it shows that a driver written for
// the reference's Simulation and Module interfaces compiles, links against
// libsphhost and runs on the GPU modules (tests/test_host.py builds it,
// tests/test_gpu_host.py runs it).
//
//   host_solver_mirror PARAMS.json IN.snap STEPS OUT.snap

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "sph/exception.hpp"
#include "sph/module.hpp"
#include "sph/simulation.hpp"
#include "sph/snapshot.hpp"
#include "sph/solver.hpp"

namespace {

class MirrorSolver {
public:
    explicit MirrorSolver(std::shared_ptr<sph::SPHParameters> param) : m_param(std::move(param)) {}

    void initialize(const std::vector<sph::SPHParticle> &initial) {
        m_sim = std::make_shared<sph::Simulation>(m_param);
        // the initial condition, as make_initial_condition leaves it
        m_sim->set_particles(initial);
        m_sim->set_particle_num(static_cast<int>(initial.size()));

        m_timestep = std::make_shared<sph::gpu::TimeStep>();
        switch (m_param->type) {
        case sph::SPHType::SSPH:
            m_pre = std::make_shared<sph::gpu::PreInteraction>();
            m_fforce = std::make_shared<sph::gpu::FluidForce>();
            break;
        case sph::SPHType::DISPH:
            m_pre = std::make_shared<sph::gpu::disph::PreInteraction>();
            m_fforce = std::make_shared<sph::gpu::disph::FluidForce>();
            break;
        case sph::SPHType::GSPH:
            m_pre = std::make_shared<sph::gpu::gsph::PreInteraction>();
            m_fforce = std::make_shared<sph::gpu::gsph::FluidForce>();
            break;
        }
        m_gforce = std::make_shared<sph::gpu::GravityForce>();

        if (m_param->type == sph::SPHType::GSPH) {
            std::vector<std::string> arrays = {"grad_density", "grad_pressure"};
            for (int k = 0; k < m_param->dim; ++k) arrays.push_back("grad_velocity_" + std::to_string(k));
            m_sim->add_vector_array(arrays);
        }
        for (auto *m : {&m_timestep, &m_pre, &m_fforce, &m_gforce}) (*m)->initialize(m_param);

        std::vector<sph::SPHParticle> &p = m_sim->get_particles();
        const int num = m_sim->get_particle_num();
        const double cs2 = m_param->physics.gamma * (m_param->physics.gamma - 1.0);
        for (int i = 0; i < num; ++i) {
            p[i].alpha = m_param->av.alpha;
            p[i].balsara = 1.0;
            p[i].sound = std::sqrt(cs2 * p[i].ene);
        }
        std::shared_ptr<sph::BHTree> tree = m_sim->get_tree();
        tree->resize(num);
        tree->make(p, num);

        m_pre->calculation(m_sim);
        m_fforce->calculation(m_sim);
        m_gforce->calculation(m_sim);
    }

    void integrate() {
        m_timestep->calculation(m_sim);
        predict();
        m_sim->make_tree();
        m_pre->calculation(m_sim);
        m_fforce->calculation(m_sim);
        m_gforce->calculation(m_sim);
        correct();
    }

    std::shared_ptr<sph::Simulation> simulation() const { return m_sim; }

private:
    // half kick of v and u, drift of x with the half-step velocity, full
    // kick of v and u, sound speed, periodic wrap
    void predict() {
        std::vector<sph::SPHParticle> &p = m_sim->get_particles();
        const sph::Periodic &periodic = *m_sim->get_periodic();
        const double dt = m_sim->get_dt();
        const double cs2 = m_param->physics.gamma * (m_param->physics.gamma - 1.0);
        for (sph::SPHParticle &q : p) {
            q.vel_p = q.vel + q.acc * (0.5 * dt);
            q.ene_p = q.ene + q.dene * (0.5 * dt);
            q.pos += q.vel_p * dt;
            q.vel += q.acc * dt;
            q.ene += q.dene * dt;
            q.sound = std::sqrt(cs2 * q.ene);
            periodic.apply(q.pos);
        }
    }

    // second half kick from the half-step state with the new derivatives
    void correct() {
        std::vector<sph::SPHParticle> &p = m_sim->get_particles();
        const double dt = m_sim->get_dt();
        const double cs2 = m_param->physics.gamma * (m_param->physics.gamma - 1.0);
        for (sph::SPHParticle &q : p) {
            q.vel = q.vel_p + q.acc * (0.5 * dt);
            q.ene = q.ene_p + q.dene * (0.5 * dt);
            q.sound = std::sqrt(cs2 * q.ene);
        }
    }

    std::shared_ptr<sph::SPHParameters> m_param;
    std::shared_ptr<sph::Simulation> m_sim;
    std::shared_ptr<sph::Module> m_timestep, m_pre, m_fforce, m_gforce;
};

}  // namespace

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: host_solver_mirror PARAMS.json IN.snap STEPS OUT.snap\n");
        return 2;
    }
    try {
        auto param = sph::read_parameter_file(argv[1]);
        const sph::Snapshot in = sph::read_snapshot(argv[2]);
        const int steps = std::atoi(argv[3]);
        MirrorSolver solver(param);
        solver.initialize(in.particles);
        auto sim = solver.simulation();
        for (int k = 0; k < steps; ++k) {
            solver.integrate();
            sim->update_time();
        }
        sph::Snapshot out;
        out.dim = param->dim;
        out.time = sim->get_time();
        out.dt = sim->get_dt();
        out.h_per_v_sig = sim->get_h_per_v_sig();
        out.particles = sim->particles_view();
        sph::write_snapshot(argv[4], out);
        std::printf("steps=%d time=%.17g\n", steps, out.time);
    } catch (const sph::SPHException &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
