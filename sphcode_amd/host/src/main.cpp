// sphgpu_run: command-line driver in the manner of the reference's main
// (src/main.cpp:6-14): read a JSON parameter file, take the initial
// condition from a snapshot, initialise and integrate on the GPU, and
// write the final state.  SPHException ends the program with status 1, as
// in the reference.
//
//   sphgpu_run PARAMS.json --print-params
//   sphgpu_run PARAMS.json --particles IN.snap [--restart] [--steps K] [--output OUT.snap] [--energy]
//              [--output-dir DIR]
//
// Without --steps the run goes to endTime (the loop of src/solver.cpp:318);
// with --output-dir it writes, as Solver::run does, DIR/energy.dat at the
// energyTime cadence and the particles at the outputTime cadence
// (full-precision snapshots, sph/output.hpp).
// With --restart IN.snap is a state written by --output after a full step:
// the run continues from its time, dt and min h/v_sig instead of
// initialising (the reference has no restart).

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "sph/exception.hpp"
#include "sph/snapshot.hpp"
#include "sph/solver.hpp"

namespace {

const char *type_name(sph::SPHType t) {
    return t == sph::SPHType::SSPH ? "ssph" : t == sph::SPHType::DISPH ? "disph" : "gsph";
}

void print_params(const sph::SPHParameters &p, const std::string &outdir) {
    std::printf("{\"outputDirectory\": \"%s\", \"DIM\": %d, \"precision\": \"%s\", \"neighborCapacity\": %d,\n",
                outdir.c_str(), p.dim, p.precision == sph::Precision::FP64 ? "fp64" : "mixed", p.neighbor_capacity);
    std::printf(" \"startTime\": %.17g, \"endTime\": %.17g, \"outputTime\": %.17g, \"energyTime\": %.17g,\n",
                p.time.start, p.time.end, p.time.output, p.time.energy);
    std::printf(" \"SPHType\": \"%s\", \"cflSound\": %.17g, \"cflForce\": %.17g,\n", type_name(p.type), p.cfl.sound,
                p.cfl.force);
    std::printf(" \"avAlpha\": %.17g, \"useBalsaraSwitch\": %s, \"useTimeDependentAV\": %s, \"alphaMax\": %.17g, "
                "\"alphaMin\": %.17g, \"epsilonAV\": %.17g,\n",
                p.av.alpha, p.av.use_balsara_switch ? "true" : "false", p.av.use_time_dependent_av ? "true" : "false",
                p.av.alpha_max, p.av.alpha_min, p.av.epsilon);
    std::printf(" \"useArtificialConductivity\": %s, \"alphaAC\": %.17g,\n", p.ac.is_valid ? "true" : "false",
                p.ac.alpha);
    std::printf(" \"maxTreeLevel\": %d, \"leafParticleNumber\": %d, \"neighborNumber\": %d, \"gamma\": %.17g,\n",
                p.tree.max_level, p.tree.leaf_particle_num, p.physics.neighbor_number, p.physics.gamma);
    std::printf(" \"kernel\": \"%s\", \"iterativeSmoothingLength\": %s, \"periodic\": %s,\n",
                p.kernel == sph::KernelType::CUBIC_SPLINE ? "cubic_spline" : "wendland",
                p.iterative_sml ? "true" : "false", p.periodic.is_valid ? "true" : "false");
    std::printf(" \"rangeMax\": [%.17g, %.17g, %.17g], \"rangeMin\": [%.17g, %.17g, %.17g],\n",
                p.periodic.range_max[0], p.periodic.range_max[1], p.periodic.range_max[2], p.periodic.range_min[0],
                p.periodic.range_min[1], p.periodic.range_min[2]);
    std::printf(" \"useGravity\": %s, \"G\": %.17g, \"theta\": %.17g, \"use2ndOrderGSPH\": %s}\n",
                p.gravity.is_valid ? "true" : "false", p.gravity.constant, p.gravity.theta,
                p.gsph.is_2nd_order ? "true" : "false");
}

void print_energy(const sph::Simulation &sim, int loop) {
    const auto e = sim.energy();
    std::printf("energy loop=%d time=%.17g kinetic=%.17g thermal=%.17g potential=%.17g total=%.17g\n", loop,
                sim.get_time(), e[0], e[1], e[2], e[3]);
}

int usage() {
    std::fprintf(stderr, "usage: sphgpu_run PARAMS.json --print-params\n"
                         "       sphgpu_run PARAMS.json --particles IN.snap [--restart] [--steps K] "
                         "[--output OUT.snap] [--energy] [--output-dir DIR]\n");
    return 2;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) return usage();
    const std::string param_file = argv[1];
    std::string particles, output, output_dir;
    int steps = -1;
    bool show = false, energy = false, restart = false;
    for (int a = 2; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--print-params")) show = true;
        else if (!std::strcmp(argv[a], "--energy")) energy = true;
        else if (!std::strcmp(argv[a], "--restart")) restart = true;
        else if (!std::strcmp(argv[a], "--particles") && a + 1 < argc) particles = argv[++a];
        else if (!std::strcmp(argv[a], "--output") && a + 1 < argc) output = argv[++a];
        else if (!std::strcmp(argv[a], "--output-dir") && a + 1 < argc) output_dir = argv[++a];
        else if (!std::strcmp(argv[a], "--steps") && a + 1 < argc) steps = std::atoi(argv[++a]);
        else return usage();
    }
    try {
        std::string outdir;
        auto param = sph::read_parameter_file(param_file, &outdir);
        if (show) {
            print_params(*param, outdir);
            return 0;
        }
        if (particles.empty()) return usage();
        sph::Snapshot in = sph::read_snapshot(particles);
        if (in.dim != param->dim) throw sph::SPHException("snapshot DIM differs from the parameter file's DIM");
        sph::Solver solver(param);
        if (restart) {
            if (!(in.h_per_v_sig > 0.0)) throw sph::SPHException("--restart needs a snapshot written after a step");
            solver.restart(in.particles, in.time, in.dt, in.h_per_v_sig);
        } else {
            solver.initialize(in.particles);
        }
        auto sim = solver.simulation();
        if (energy) print_energy(*sim, 0);
        int loop = 0;
        if (steps < 0) {
            std::unique_ptr<sph::Output> out;
            if (!output_dir.empty()) out.reset(new sph::Output(output_dir));
            loop = solver.run(out.get());
        } else {
            for (; loop < steps; ++loop) {
                solver.integrate();
                if (energy) print_energy(*sim, loop + 1);
            }
        }
        std::printf("steps=%d time=%.17g\n", loop, sim->get_time());
        if (!output.empty()) {
            sph::Snapshot out;
            out.dim = param->dim;
            out.time = sim->get_time();
            out.dt = sim->get_dt();
            out.h_per_v_sig = sim->get_h_per_v_sig();
            out.particles = sim->particles_view();
            sph::write_snapshot(output, out);
        }
    } catch (const sph::SPHException &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
