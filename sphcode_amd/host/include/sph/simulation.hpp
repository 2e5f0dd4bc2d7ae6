// Simulation state seen by the modules (the reference's Simulation,
// include/simulation.hpp:24-45).  The particles live on the GPU; the host
// vector is a copy refreshed on request (sync_from_device) for output and
// tests, and pushed with set_particles.
#pragma once

#include <array>
#include <memory>
#include <vector>

#include "sph/parameters.hpp"

namespace sph {

// One particle (include/particle.hpp:8-33); vectors hold DIM components.
struct SPHParticle {
    double pos[3] = {0, 0, 0};
    double vel[3] = {0, 0, 0};
    double vel_p[3] = {0, 0, 0};
    double acc[3] = {0, 0, 0};
    double mass = 0, dens = 0, pres = 0, ene = 0, ene_p = 0, dene = 0, sml = 0, sound = 0;
    double balsara = 0, alpha = 0, gradh = 0, phi = 0;
    int id = 0;
    int neighbor = 0;
};

class Simulation {
public:
    explicit Simulation(std::shared_ptr<SPHParameters> param);
    ~Simulation();
    Simulation(const Simulation &) = delete;
    Simulation &operator=(const Simulation &) = delete;

    void set_particles(const std::vector<SPHParticle> &particles);
    std::vector<SPHParticle> &get_particles() { return m_particles; }
    void sync_from_device();
    int get_particle_num() const { return (int)m_particles.size(); }

    double get_time() const;
    double get_dt() const;
    double get_h_per_v_sig() const;
    void set_dt(double dt);
    void update_time();  // time += dt (src/simulation.cpp:32-35)
    // continue from a state saved after a full step: time, dt and min h/v_sig
    // as saved, no initial smoothing, tree rebuilt (sphgpu_restart)
    void restart(double time, double dt, double h_per_v_sig);

    void make_tree();    // BHTree::make on the device (src/simulation.cpp:37-40)

    // kinetic, thermal, potential, total, momentum[DIM] (src/output.cpp:66-90)
    std::array<double, 7> energy() const;

    const SPHParameters &params() const { return *m_param; }
    void *device() const { return m_ctx; }

private:
    std::shared_ptr<SPHParameters> m_param;
    std::vector<SPHParticle> m_particles;
    void *m_ctx = nullptr;
};

}  // namespace sph
