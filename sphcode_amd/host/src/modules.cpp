// GPU Modules: each calculation() is one stage of the C ABI on the
// simulation's device state.

#include "sph/exception.hpp"
#include "sph/module.hpp"

namespace sph {
namespace gpu {

namespace {
void need(const std::shared_ptr<Simulation> &sim) {
    if (!sim) throw SPHException("calculation: no simulation");
}
const char *type_name(SPHType t) { return t == SPHType::SSPH ? "ssph" : t == SPHType::DISPH ? "disph" : "gsph"; }
void check_type(SPHType have, SPHType want, const char *module) {
    if (have != want)
        throw SPHException(std::string(module) + ": module is for " + type_name(want) + " but SPHType is " +
                           type_name(have));
}
}  // namespace

// src/timestep.cpp:18-38
void TimeStep::initialize(std::shared_ptr<SPHParameters>) {}
void TimeStep::calculation(std::shared_ptr<Simulation> sim) {
    need(sim);
    double dt = 0.0;
    check_status(sphgpu_timestep(sim->device(), &dt), "sphgpu_timestep");
}

// src/pre_interaction.cpp:19-37 (parameters are taken from the simulation's
// context, created from the same SPHParameters)
void PreInteraction::initialize(std::shared_ptr<SPHParameters> param) {
    if (!param) throw SPHException("PreInteraction::initialize: no parameters");
    check_type(param->type, expected(), "PreInteraction");
}
void PreInteraction::calculation(std::shared_ptr<Simulation> sim) {
    need(sim);
    check_type(sim->params().type, expected(), "PreInteraction");
    check_status(sphgpu_pre_interaction(sim->device()), "sphgpu_pre_interaction");
    sim->note_pre_interaction();
}

void FluidForce::initialize(std::shared_ptr<SPHParameters> param) {
    if (!param) throw SPHException("FluidForce::initialize: no parameters");
    check_type(param->type, expected(), "FluidForce");
}
void FluidForce::calculation(std::shared_ptr<Simulation> sim) {
    need(sim);
    check_type(sim->params().type, expected(), "FluidForce");
    check_status(sphgpu_fluid_force(sim->device()), "sphgpu_fluid_force");
}

// src/gravity_force.cpp:44-58: without gravity the stage does nothing
void GravityForce::initialize(std::shared_ptr<SPHParameters> param) {
    if (!param) throw SPHException("GravityForce::initialize: no parameters");
    m_is_valid = param->gravity.is_valid;
}
void GravityForce::calculation(std::shared_ptr<Simulation> sim) {
    need(sim);
    if (!m_is_valid) return;
    check_status(sphgpu_gravity_force(sim->device()), "sphgpu_gravity_force");
}

}  // namespace gpu
}  // namespace sph
