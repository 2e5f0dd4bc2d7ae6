// Stage parameters, field for field those of the reference's SPHParameters
// (include/parameters.hpp:8-79), with the compile-time DIM of
// include/defines.hpp:5 and this build's precision and list-capacity
// settings as run-time members.
#pragma once

#include "sphgpu.h"

namespace sph {

enum struct SPHType { SSPH, DISPH, GSPH };
enum struct KernelType { CUBIC_SPLINE, WENDLAND, UNKNOWN };
enum struct Precision { FP64, MIXED };

struct SPHParameters {
    struct Time {
        double start = 0.0, end = 0.0, output = 0.0, energy = 0.0;
    } time;
    SPHType type = SPHType::SSPH;
    struct CFL {
        double sound = 0.3, force = 0.125;
    } cfl;
    struct ArtificialViscosity {
        double alpha = 1.0;
        bool use_balsara_switch = true;
        bool use_time_dependent_av = false;
        double alpha_max = 2.0, alpha_min = 0.1;
        double epsilon = 0.2;  // tau = h / (epsilon c)
    } av;
    struct ArtificialConductivity {
        double alpha = 1.0;
        bool is_valid = false;
    } ac;
    struct Tree {
        int max_level = 20, leaf_particle_num = 1;
    } tree;
    struct Physics {
        int neighbor_number = 32;
        double gamma = 0.0;
    } physics;
    KernelType kernel = KernelType::CUBIC_SPLINE;
    bool iterative_sml = true;
    struct Periodic {
        bool is_valid = false;
        double range_max[3] = {0, 0, 0}, range_min[3] = {0, 0, 0};
    } periodic;
    struct Gravity {
        bool is_valid = false;
        double constant = 1.0, theta = 0.5;
    } gravity;
    struct GSPH {
        bool is_2nd_order = true;
    } gsph;

    // build-specific
    int dim = 3;
    Precision precision = Precision::FP64;
    int neighbor_capacity = 0;  // 0: neighbor_number * 20
};

// The C ABI view of the parameters.
sphgpu_params to_c_params(const SPHParameters &p);

}  // namespace sph
