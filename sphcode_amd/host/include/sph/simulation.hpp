// Simulation state seen by the modules, with the accessors of the
// reference's Simulation (include/simulation.hpp:24-45, src/simulation.cpp):
// particles, particle number, time, dt, min h/v_sig, kernel, periodic
// boundary, tree, and named scalar / vector arrays.
//
// The particles live on the GPU (a sphgpu context).  get_particles() hands
// out a host copy that is refreshed from the device when the device is
// newer, and is taken to be modified by the caller: the next device
// operation (a module's calculation, make_tree, the tree's make) uploads
// it first.  So host code written against the reference -- e.g. a Solver
// whose predict / correct loop over get_particles() (src/solver.cpp:431-474)
// -- sees and changes the device state, at the cost of a transfer each way.
// The GPU Solver (sph/solver.hpp) never touches the host copy.
//
// GSPH's gradient arrays (add_vector_array with "grad_density",
// "grad_pressure", "grad_velocity_0" .. "_<DIM-1>", src/solver.cpp:373-385)
// travel with the particles: grad_velocity_k[i][l] = d v_k / d x_l.
#pragma once

#include <array>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "sph/exception.hpp"
#include "sph/kernel_function.hpp"
#include "sph/parameters.hpp"
#include "sph/particle.hpp"
#include "sph/periodic.hpp"

namespace sph {

class Simulation;

// The tree handle of Simulation::get_tree().  The octree is built and kept
// on the device; the reference's BHTree::resize (node pool for num
// particles, include/bhtree.hpp:77) and make (src/bhtree.cpp:55-107) map
// to the context.
class BHTree {
public:
    explicit BHTree(Simulation &sim) : m_sim(sim) {}
    // the node pool is sized with the particles (5 N + 1 nodes); num must be
    // the particle number
    void resize(int num);
    // builds the device tree from the particles (uploading them first when
    // the host copy was changed); num must be the particle number
    void make(std::vector<SPHParticle> &particles, int num);

private:
    Simulation &m_sim;
};

class Simulation {
public:
    explicit Simulation(std::shared_ptr<SPHParameters> param);
    ~Simulation();
    Simulation(const Simulation &) = delete;
    Simulation &operator=(const Simulation &) = delete;

    // --- the reference's accessors
    // host copy of the particles (caller order), current with the device;
    // treated as modified by the caller
    std::vector<SPHParticle> &get_particles();
    void set_particles(const std::vector<SPHParticle> &particles);
    int get_particle_num() const { return m_particle_num; }
    void set_particle_num(int num);
    double get_time() const;
    void set_time(double time);
    double get_dt() const;
    void set_dt(double dt);
    double get_h_per_v_sig() const;
    // marks the state as being in mid-run (no initial smoothing follows)
    void set_h_per_v_sig(double h_per_v_sig);
    std::shared_ptr<KernelFunction> get_kernel() const { return m_kernel; }
    std::shared_ptr<Periodic> get_periodic() const { return m_periodic; }
    std::shared_ptr<BHTree> get_tree() { return m_tree; }
    void update_time();  // time += dt (src/simulation.cpp:32-35)
    void make_tree();    // BHTree::make on the device (src/simulation.cpp:37-40)
    void add_scalar_array(const std::vector<std::string> &names);
    void add_vector_array(const std::vector<std::string> &names);
    std::vector<real> &get_scalar_array(const std::string &name);
    std::vector<vec_t> &get_vector_array(const std::string &name);

    // --- this build's
    // the current particles for reading, without marking them modified
    const std::vector<SPHParticle> &particles_view();
    void sync_from_device();  // refresh the host copy now
    // continue from a state saved after a full step: time, dt and min
    // h/v_sig as saved, no initial smoothing, tree rebuilt (sphgpu_restart)
    void restart(double time, double dt, double h_per_v_sig);
    // kinetic, thermal, potential, total, momentum[DIM] (src/output.cpp:66-90)
    std::array<double, 7> energy() const;
    const SPHParameters &params() const { return *m_param; }
    // the device context for a stage call: pending host changes are
    // uploaded first, and the host copy is stale afterwards
    void *device();
    // called by the pre-interaction module: from here on an upload of host
    // changes continues the run instead of starting one
    void note_pre_interaction() { m_started = true; }

private:
    void push();
    void pull();

    std::shared_ptr<SPHParameters> m_param;
    std::vector<SPHParticle> m_particles;
    int m_particle_num = 0;
    void *m_ctx = nullptr;
    bool m_host_current = true;   // host copy equals the device state
    bool m_host_dirty = false;    // host copy may differ from the device
    bool m_started = false;
    std::shared_ptr<KernelFunction> m_kernel;
    std::shared_ptr<Periodic> m_periodic;
    std::shared_ptr<BHTree> m_tree;
    std::map<std::string, std::vector<real>> m_scalar_arrays;
    std::map<std::string, std::vector<vec_t>> m_vector_arrays;
};

}  // namespace sph
