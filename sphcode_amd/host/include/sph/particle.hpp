// SPHParticle, field for field the reference's (include/particle.hpp:8-33).
#pragma once

#include "sph/vector_type.hpp"

namespace sph {

struct SPHParticle {
    vec_t pos;    // position
    vec_t vel;    // velocity
    vec_t vel_p;  // velocity at t + dt/2
    vec_t acc;    // acceleration
    real mass = 0.0;
    real dens = 0.0;
    real pres = 0.0;
    real ene = 0.0;    // specific internal energy
    real ene_p = 0.0;  // ene at t + dt/2
    real dene = 0.0;   // du/dt
    real sml = 0.0;    // smoothing length
    real sound = 0.0;
    real balsara = 0.0;
    real alpha = 0.0;
    real gradh = 0.0;
    real phi = 0.0;
    int id = 0;
    int neighbor = 0;
    // the reference's tree threads particles through this pointer; the
    // device tree does not use it
    SPHParticle *next = nullptr;
};

}  // namespace sph
