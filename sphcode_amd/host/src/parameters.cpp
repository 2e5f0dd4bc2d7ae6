// JSON parameter reader with the reference's keys, defaults and checks
// (src/solver.cpp:155-299), and the conversion to the C ABI struct.

#include <fstream>
#include <sstream>

#include "json.hpp"
#include "sph/exception.hpp"
#include "sph/parameters.hpp"
#include "sph/solver.hpp"

namespace sph {

sphgpu_params to_c_params(const SPHParameters &p) {
    sphgpu_params c;
    check_status(sphgpu_params_default(&c), "sphgpu_params_default");
    c.dim = p.dim;
    c.sph_type = p.type == SPHType::SSPH ? SPHGPU_SSPH : p.type == SPHType::DISPH ? SPHGPU_DISPH : SPHGPU_GSPH;
    if (p.kernel == KernelType::UNKNOWN) throw SPHException("kernel is unknown.");
    c.kernel = p.kernel == KernelType::CUBIC_SPLINE ? SPHGPU_CUBIC_SPLINE : SPHGPU_WENDLAND;
    c.precision = p.precision == Precision::FP64 ? SPHGPU_FP64 : SPHGPU_MIXED;
    c.time_start = p.time.start;
    c.time_end = p.time.end;
    c.time_output = p.time.output;
    c.time_energy = p.time.energy;
    c.cfl_sound = p.cfl.sound;
    c.cfl_force = p.cfl.force;
    c.av_alpha = p.av.alpha;
    c.av_alpha_max = p.av.alpha_max;
    c.av_alpha_min = p.av.alpha_min;
    c.av_epsilon = p.av.epsilon;
    c.av_use_balsara_switch = p.av.use_balsara_switch;
    c.av_use_time_dependent = p.av.use_time_dependent_av;
    c.ac_alpha = p.ac.alpha;
    c.ac_is_valid = p.ac.is_valid;
    c.tree_max_level = p.tree.max_level;
    c.tree_leaf_particle_num = p.tree.leaf_particle_num;
    c.neighbor_number = p.physics.neighbor_number;
    c.gamma = p.physics.gamma;
    c.iterative_sml = p.iterative_sml;
    c.periodic_is_valid = p.periodic.is_valid;
    for (int k = 0; k < 3; ++k) {
        c.range_max[k] = p.periodic.range_max[k];
        c.range_min[k] = p.periodic.range_min[k];
    }
    c.gravity_is_valid = p.gravity.is_valid;
    c.gravity_constant = p.gravity.constant;
    c.gravity_theta = p.gravity.theta;
    c.gsph_is_2nd_order = p.gsph.is_2nd_order;
    c.neighbor_capacity = p.neighbor_capacity;
    return c;
}

std::shared_ptr<SPHParameters> read_parameters(const std::string &text, std::string *output_directory) {
    using namespace json;
    const Value in = parse(text);
    if (in.kind != Value::OBJECT) throw SPHException("JSON parse error: the document is not an object");
    auto p = std::make_shared<SPHParameters>();

    // build-specific keys
    p->dim = get_int(in, "DIM", 3);
    if (p->dim < 1 || p->dim > 3) throw SPHException("DIM must be 1, 2 or 3");
    const std::string prec = get_string(in, "precision", "fp64");
    if (prec == "fp64") p->precision = Precision::FP64;
    else if (prec == "mixed") p->precision = Precision::MIXED;
    else throw SPHException("Unknown precision");
    p->neighbor_capacity = get_int(in, "neighborCapacity", 0);

    const std::string outdir = get_string(in, "outputDirectory");
    if (output_directory) *output_directory = outdir;

    // time
    p->time.start = get_real(in, "startTime", 0.0);
    p->time.end = get_real(in, "endTime");
    if (p->time.end < p->time.start) throw SPHException("endTime < startTime");
    p->time.output = get_real(in, "outputTime", (p->time.end - p->time.start) / 100);
    p->time.energy = get_real(in, "energyTime", p->time.output);

    // type
    const std::string type = get_string(in, "SPHType", "ssph");
    if (type == "ssph") p->type = SPHType::SSPH;
    else if (type == "disph") p->type = SPHType::DISPH;
    else if (type == "gsph") p->type = SPHType::GSPH;
    else throw SPHException("Unknown SPH type");

    // CFL
    p->cfl.sound = get_real(in, "cflSound", 0.3);
    p->cfl.force = get_real(in, "cflForce", 0.125);

    // artificial viscosity
    p->av.alpha = get_real(in, "avAlpha", 1.0);
    p->av.use_balsara_switch = get_bool(in, "useBalsaraSwitch", true);
    p->av.use_time_dependent_av = get_bool(in, "useTimeDependentAV", false);
    if (p->av.use_time_dependent_av) {
        p->av.alpha_max = get_real(in, "alphaMax", 2.0);
        p->av.alpha_min = get_real(in, "alphaMin", 0.1);
        if (p->av.alpha_max < p->av.alpha_min) throw SPHException("alphaMax < alphaMin");
        p->av.epsilon = get_real(in, "epsilonAV", 0.2);
    }

    // artificial conductivity
    p->ac.is_valid = get_bool(in, "useArtificialConductivity", false);
    if (p->ac.is_valid) p->ac.alpha = get_real(in, "alphaAC", 1.0);

    // tree
    p->tree.max_level = get_int(in, "maxTreeLevel", 20);
    p->tree.leaf_particle_num = get_int(in, "leafParticleNumber", 1);

    // physics
    p->physics.neighbor_number = get_int(in, "neighborNumber", 32);
    p->physics.gamma = get_real(in, "gamma");

    // kernel
    const std::string kernel = get_string(in, "kernel", "cubic_spline");
    if (kernel == "cubic_spline") p->kernel = KernelType::CUBIC_SPLINE;
    else if (kernel == "wendland") p->kernel = KernelType::WENDLAND;
    else throw SPHException("kernel is unknown.");
    // C4Kernel asserts DIM != 1 (include/kernel/wendland_kernel.hpp:25-28)
    if (p->kernel == KernelType::WENDLAND && p->dim == 1)
        throw SPHException("Wendland C4 kernel is not defined for DIM = 1");

    p->iterative_sml = get_bool(in, "iterativeSmoothingLength", true);

    // periodic
    p->periodic.is_valid = get_bool(in, "periodic", false);
    if (p->periodic.is_valid) {
        const auto rmax = get_real_array(in, "rangeMax");
        if ((int)rmax.size() != p->dim) throw SPHException("rangeMax != DIM");
        const auto rmin = get_real_array(in, "rangeMin");
        // the reference reports this case with the rangeMax message too
        if ((int)rmin.size() != p->dim) throw SPHException("rangeMax != DIM");
        for (int k = 0; k < p->dim; ++k) {
            p->periodic.range_max[k] = rmax[k];
            p->periodic.range_min[k] = rmin[k];
        }
    }

    // gravity
    p->gravity.is_valid = get_bool(in, "useGravity", false);
    if (p->gravity.is_valid) {
        p->gravity.constant = get_real(in, "G", 1.0);
        p->gravity.theta = get_real(in, "theta", 0.5);
    }

    // GSPH
    if (p->type == SPHType::GSPH) p->gsph.is_2nd_order = get_bool(in, "use2ndOrderGSPH", true);

    return p;
}

std::shared_ptr<SPHParameters> read_parameter_file(const std::string &path, std::string *output_directory) {
    std::ifstream f(path);
    if (!f) throw SPHException(path + ": cannot open file");
    std::stringstream ss;
    ss << f.rdbuf();
    return read_parameters(ss.str(), output_directory);
}

}  // namespace sph
