// The stage interface of the reference (include/module.hpp:10-14) and its
// GPU implementations.  Each calculation() is one call of the C ABI.
#pragma once

#include <memory>

#include "sph/parameters.hpp"
#include "sph/simulation.hpp"

namespace sph {

class Module {
public:
    virtual ~Module() = default;
    virtual void initialize(std::shared_ptr<SPHParameters> param) = 0;
    virtual void calculation(std::shared_ptr<Simulation> sim) = 0;
};

namespace gpu {

// TimeStep (include/timestep.hpp:8-14): dt from the device reductions.
class TimeStep : public Module {
public:
    void initialize(std::shared_ptr<SPHParameters> param) override;
    void calculation(std::shared_ptr<Simulation> sim) override;
};

// PreInteraction (include/pre_interaction.hpp:13-39) for the SPH type the
// simulation was created with; the disph/gsph classes below only check
// that the type matches, as the reference selects them by SPHType
// (src/solver.cpp:359-370).
class PreInteraction : public Module {
public:
    void initialize(std::shared_ptr<SPHParameters> param) override;
    void calculation(std::shared_ptr<Simulation> sim) override;

protected:
    virtual SPHType expected() const { return SPHType::SSPH; }
};

class FluidForce : public Module {
public:
    void initialize(std::shared_ptr<SPHParameters> param) override;
    void calculation(std::shared_ptr<Simulation> sim) override;

protected:
    virtual SPHType expected() const { return SPHType::SSPH; }
};

namespace disph {
class PreInteraction : public gpu::PreInteraction {
protected:
    SPHType expected() const override { return SPHType::DISPH; }
};
class FluidForce : public gpu::FluidForce {
protected:
    SPHType expected() const override { return SPHType::DISPH; }
};
}  // namespace disph

namespace gsph {
class PreInteraction : public gpu::PreInteraction {
protected:
    SPHType expected() const override { return SPHType::GSPH; }
};
class FluidForce : public gpu::FluidForce {
protected:
    SPHType expected() const override { return SPHType::GSPH; }
};
}  // namespace gsph

// GravityForce (include/gravity_force.hpp:10-17).
class GravityForce : public Module {
public:
    void initialize(std::shared_ptr<SPHParameters> param) override;
    void calculation(std::shared_ptr<Simulation> sim) override;

private:
    bool m_is_valid = false;
};

}  // namespace gpu
}  // namespace sph
