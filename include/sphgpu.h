/*
 * sphgpu.h -- C ABI of the MI355X (gfx950) SPH per-step particle pipeline.
 *
 * The reference (mitchiinaga/sphcode) drives one time step through four
 * Module objects plus a tree rebuild and two kick-drift loops:
 *
 *   Solver::integrate                       src/solver.cpp:417-429
 *     TimeStep::calculation                 src/timestep.cpp:18-38
 *     Solver::predict                       src/solver.cpp:431-456
 *     Simulation::make_tree -> BHTree::make src/simulation.cpp:37-40, src/bhtree.cpp:55-107
 *     PreInteraction::calculation           src/pre_interaction.cpp:39-169
 *         (disph: src/disph/d_pre_interaction.cpp:21-162, gsph: src/gsph/g_pre_interaction.cpp:27-144)
 *     FluidForce::calculation               src/fluid_force.cpp:26-86
 *         (disph: src/disph/d_fluid_force.cpp:26-86, gsph: src/gsph/g_fluid_force.cpp:39-166)
 *     GravityForce::calculation             src/gravity_force.cpp:52-89
 *     Solver::correct                       src/solver.cpp:458-474
 *
 * Every entry point below replaces one of those calls.  The Module
 * interface (include/module.hpp:10-14) is kept on the C++ host side
 * (sphcode_amd/host), whose Module subclasses call these functions.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative SPHGPU_E* code on
 *     failure; sphgpu_last_error() returns a message for the calling thread.
 *     (The reference throws SPHException, include/exception.hpp:9-14.)
 *   - Device buffers belong to the context; host buffers belong to the
 *     caller.  Work is ordered on the context's HIP stream; functions that
 *     return host values synchronise that stream.
 *   - One host thread per context.  No torch types cross this boundary.
 *   - Host particle arrays are in the caller's particle order (the order
 *     given to sphgpu_upload, i.e. SPHParticle::id); vectors are stored
 *     [i*dim + k].
 */
#ifndef SPHGPU_H
#define SPHGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes. */
#define SPHGPU_OK 0
#define SPHGPU_E_INVALID (-1)        /* bad argument or parameter           */
#define SPHGPU_E_HIP (-2)            /* HIP runtime error                   */
#define SPHGPU_E_NO_DEVICE (-3)      /* no usable gfx950 device             */
#define SPHGPU_E_LIST_OVERFLOW (-4)  /* a neighbour list exceeded capacity  */
#define SPHGPU_E_TREE (-5)           /* tree build failed (node pool)       */
#define SPHGPU_E_STATE (-6)          /* call out of order (e.g. no upload)  */

/* SPHType (include/parameters.hpp:8-12). */
#define SPHGPU_SSPH 0
#define SPHGPU_DISPH 1
#define SPHGPU_GSPH 2

/* KernelType (include/parameters.hpp:14-18). */
#define SPHGPU_CUBIC_SPLINE 0
#define SPHGPU_WENDLAND 1

/* Arithmetic of the pair sums. */
#define SPHGPU_FP64 0   /* all FP64, as real=double (include/defines.hpp:6)   */
#define SPHGPU_MIXED 1  /* FP32 pair arithmetic relative to particle i, FP32
                           terms summed in FP64; FP64 kick-drift, Newton
                           update and test, and global reductions            */

/*
 * sphgpu_params mirrors SPHParameters (include/parameters.hpp:20-79) field by
 * field, plus the compile-time DIM (include/defines.hpp:5) and the build's
 * own precision / capacity settings.  Defaults of the reference's JSON reader
 * (src/solver.cpp:155-299) are applied by sphgpu_params_default().
 */
typedef struct sphgpu_params {
    int32_t dim;                    /* DIM: 1, 2 or 3                        */
    int32_t sph_type;               /* SPHGPU_SSPH / _DISPH / _GSPH          */
    int32_t kernel;                 /* SPHGPU_CUBIC_SPLINE / _WENDLAND       */
    int32_t precision;              /* SPHGPU_FP64 / SPHGPU_MIXED            */

    double time_start;              /* time.start                            */
    double time_end;                /* time.end                              */
    double time_output;             /* time.output                           */
    double time_energy;             /* time.energy                           */

    double cfl_sound;               /* cfl.sound                             */
    double cfl_force;               /* cfl.force                             */

    double av_alpha;                /* av.alpha                              */
    double av_alpha_max;            /* av.alpha_max                          */
    double av_alpha_min;            /* av.alpha_min                          */
    double av_epsilon;              /* av.epsilon                            */
    int32_t av_use_balsara_switch;  /* av.use_balsara_switch                 */
    int32_t av_use_time_dependent;  /* av.use_time_dependent_av              */

    double ac_alpha;                /* ac.alpha                              */
    int32_t ac_is_valid;            /* ac.is_valid                           */

    int32_t tree_max_level;         /* tree.max_level                        */
    int32_t tree_leaf_particle_num; /* tree.leaf_particle_num                */

    int32_t neighbor_number;        /* physics.neighbor_number               */
    double gamma;                   /* physics.gamma                         */

    int32_t iterative_sml;          /* iterative_sml                         */

    int32_t periodic_is_valid;      /* periodic.is_valid                     */
    double range_max[3];            /* periodic.range_max (first dim used)   */
    double range_min[3];            /* periodic.range_min                    */

    int32_t gravity_is_valid;       /* gravity.is_valid                      */
    double gravity_constant;        /* gravity.constant                      */
    double gravity_theta;           /* gravity.theta                         */

    int32_t gsph_is_2nd_order;      /* gsph.is_2nd_order                     */

    /* Per-particle neighbour-list capacity.  0 selects
       neighbor_number * 20, the reference's capacity
       (include/defines.hpp:29).  Exceeding it is reported as
       SPHGPU_E_LIST_OVERFLOW instead of the reference's unchecked write
       (src/bhtree.cpp:257-258). */
    int32_t neighbor_capacity;
    int32_t reserved0;
} sphgpu_params;

/*
 * Structure-of-arrays view of SPHParticle (include/particle.hpp:8-33) in the
 * caller's memory.  Any pointer may be NULL: on upload the field is then
 * zero-initialised (id defaults to the array index), on download it is
 * skipped.  grad_* are the GSPH extra arrays (src/solver.cpp:373-385):
 * grad_density and grad_pressure are n*dim, grad_velocity is n*dim*dim with
 * grad_velocity[(i*dim + k)*dim + l] = d v_k / d x_l.
 */
typedef struct sphgpu_particles {
    int64_t n;
    double *pos, *vel, *vel_p, *acc;
    double *mass, *dens, *pres, *ene, *ene_p, *dene, *sml, *sound;
    double *balsara, *alpha, *gradh, *phi;
    int32_t *id, *neighbor;
    double *grad_density, *grad_pressure, *grad_velocity;
} sphgpu_particles;

/* Counters accumulated since the last sphgpu_reset_stats(). */
typedef struct sphgpu_stats {
    int64_t steps;                /* full integrate() steps                   */
    int64_t gather_pairs;         /* pairs with r < h_i in PreInteraction     */
    int64_t force_pairs;          /* pairs with r < max(h_i,h_j), r != 0      */
    int64_t gravity_pp;           /* particle-particle gravity interactions   */
    int64_t gravity_pn;           /* particle-node (monopole) interactions    */
    int64_t newton_failures;      /* non-converged Newton solves
                                     (src/pre_interaction.cpp:277-282)        */
    int64_t newton_iterations;    /* total Newton iterations                  */
    int64_t search_truncations;   /* Newton h beyond the search radius        */
    int64_t list_overflows;       /* neighbour-list capacity exceeded         */
    int64_t max_list_length;      /* longest neighbour list seen              */
    int64_t tree_nodes;           /* nodes of the last tree                   */
    int64_t tree_max_level;       /* deepest level of the last tree           */
    int64_t gravity_opened;       /* cells opened by the gravity walk         */
    int64_t search_nodes;         /* cells visited by neighbour-search walks  */
    int64_t search_candidates;    /* particles tested by neighbour searches   */
    int64_t gravity_wave_steps;   /* iterations of the gravity walk, summed over
                                     waves (each visits one node per iteration
                                     for the lanes whose traversal is there)   */
    int64_t gravity_rewalks;      /* 64-particle groups whose record of opened
                                     leaves overflowed, walked again in full   */
    int64_t gravity_pp_slots;     /* lane slots of the particle-pair pass: 64
                                     per iteration of a 64-lane group's pair
                                     loop (as many as its lane with the most
                                     pairs has; for a group walked again in
                                     full, per particle of each leaf that one
                                     of its lanes opened); gravity_pp / this
                                     is the pass's lane efficiency             */
    int64_t gravity_max_records;  /* most leaves one 64-particle group opened
                                     (records above 1024 per group fall back
                                     to the full walk)                         */
} sphgpu_stats;

/* Stage timers (milliseconds, HIP events on the context stream),
   accumulated over sphgpu_integrate calls since the last
   sphgpu_reset_stats().  The last five are spans of single kernels (with
   their small record-packing launches) inside those stages:
     search        the neighbour search of the pre-interaction stage
     pre_pairs     the pre-interaction's pair pass (Newton h, sums)
     force_pairs   the fluid force's pair pass
     gravity_walk  the gravity tree walk (on the gravity stream)
     gravity_leaf  the gravity particle-pair pass (on the gravity stream) */
typedef struct sphgpu_timers {
    double timestep, predict, tree, pre_interaction, fluid_force, gravity, correct;
    double search, pre_pairs, force_pairs, gravity_walk, gravity_leaf;
} sphgpu_timers;

/* ---------------------------------------------------------------------- */

const char *sphgpu_last_error(void);
int sphgpu_version(void);

/* Fill *p with the reference's defaults (src/solver.cpp:155-299); the
   required keys (gamma, endTime) are set to 0 and must be supplied. */
int sphgpu_params_default(sphgpu_params *p);

/* Simulation(param) + module initialize(param)
   (src/simulation.cpp:12-30, src/solver.cpp:387-390). */
int sphgpu_create(const sphgpu_params *params, void **ctx);
int sphgpu_destroy(void *ctx);

/* Set / get the particle array (Simulation::set_particles).  Upload resets
   the state so that the next pre-interaction performs the reference's
   one-off initial smoothing (src/pre_interaction.cpp:41-44). */
int sphgpu_upload(void *ctx, const sphgpu_particles *host);
int sphgpu_download(void *ctx, sphgpu_particles *host);

/* Solver::initialize after the initial condition (src/solver.cpp:392-414):
   alpha = avAlpha, balsara = 1, c = sqrt(gamma (gamma-1) u), tree build,
   then PreInteraction, FluidForce and GravityForce once. */
int sphgpu_initialize(void *ctx);

/* The individual stages of one step, in Solver::integrate order. */
int sphgpu_timestep(void *ctx, double *dt_out);   /* TimeStep::calculation  */
int sphgpu_predict(void *ctx);                    /* Solver::predict        */
int sphgpu_make_tree(void *ctx);                  /* Simulation::make_tree  */
int sphgpu_pre_interaction(void *ctx);            /* PreInteraction         */
int sphgpu_fluid_force(void *ctx);                /* FluidForce             */
int sphgpu_gravity_force(void *ctx);              /* GravityForce           */
int sphgpu_correct(void *ctx);                    /* Solver::correct        */

/* Solver::integrate followed by Simulation::update_time
   (src/solver.cpp:417-429, src/simulation.cpp:32-35).  If fixed_dt > 0
   the TimeStep stage is replaced by dt = fixed_dt. */
int sphgpu_integrate(void *ctx, double fixed_dt);

/* Simulation scalars. */
int sphgpu_get_time(void *ctx, double *time, double *dt, double *h_per_v_sig);
int sphgpu_set_time(void *ctx, double time, double dt);

/* Resume a run from a state saved after a full step (the reference has no
   restart; its output is lossy text, src/output.cpp:14-64).  After
   sphgpu_upload of the saved particles (h, density, alpha, acceleration and
   du/dt included), this sets the time, dt and min h/v_sig the next
   sphgpu_integrate uses, skips the one-off initial smoothing and builds the
   tree.  A continued run then matches the uninterrupted one to round-off
   (particles in the same octree cell may be summed in another order). */
int sphgpu_restart(void *ctx, double time, double dt, double h_per_v_sig);

/* Global FP64 reductions of the reference's energy output
   (src/output.cpp:66-90): out[0]=kinetic, out[1]=thermal,
   out[2]=potential (sum of m phi / 2), out[3]=total, out[4..4+dim) =
   total momentum. */
int sphgpu_energy(void *ctx, double *out);

int sphgpu_get_stats(void *ctx, sphgpu_stats *stats);
int sphgpu_get_timers(void *ctx, sphgpu_timers *timers);
int sphgpu_reset_stats(void *ctx);
int sphgpu_enable_timers(void *ctx, int enable);

/* Diagnostic counters (no counterpart in the reference): with enable != 0
   the kernels count the interactions, Newton iterations and node visits
   that sphgpu_get_stats reports (gather_pairs, force_pairs, gravity_pp,
   gravity_pn, gravity_opened, gravity_wave_steps, gravity_pp_slots,
   gravity_max_records, newton_iterations, search_nodes,
   search_candidates, max_list_length).  Off by default: the gravity walk
   spends scalar instructions on every node visit to count.  The events
   that change the result or report an error (truncations, list overflows,
   Newton failures, tree errors, gravity rewalks) are always counted. */
int sphgpu_enable_stats(void *ctx, int enable);

/* Neighbour sets of the last pre-interaction (kind 0: gather candidates,
   r^2 < (kernel_ratio*h_guess)^2, src/bhtree.cpp:234-270) or the last fluid
   force (kind 1: r < max(h_i,h_j), r != 0, src/fluid_force.cpp:62-64),
   available when params were created with record_sets enabled through
   sphgpu_set_record_sets().  counts[i] receives the set size of particle i
   (caller order); lists receives the member ids, ascending, packed
   back-to-back in caller order, up to cap entries in total. */
int sphgpu_set_record_sets(void *ctx, int enable);
int sphgpu_neighbor_sets(void *ctx, int kind, int32_t *counts, int32_t *lists, int64_t cap);

/* Tree of the last make_tree (debug/parity): number of nodes, and for each
   node in depth-first pre-order its level, particle count, edge, centre,
   mass and centre of mass (host arrays sized by the caller; NULL skips). */
int sphgpu_tree_info(void *ctx, int64_t *num_nodes, int32_t *level, int32_t *count,
                     double *edge, double *center, double *mass, double *com,
                     int64_t cap);

int sphgpu_synchronize(void *ctx);

/*
 * Differentiable rollout.  No counterpart in the reference: this is the
 * gradient of a K-step trajectory with respect to its initial state.
 *
 * sphgpu_rollout runs Solver::initialize on the uploaded state and then
 * `steps` integrate() steps with dt fixed (as sphgpu_integrate(ctx, dt)),
 * keeping a checkpoint of the state at the start of every step.  The
 * context ends at the final state.
 *
 * sphgpu_rollout_adjoint takes the gradient of a scalar loss with respect
 * to the final positions, velocities and specific energies (caller order,
 * n*dim / n*dim / n doubles; NULL = zero) and returns its gradient with
 * respect to the initial ones (NULL outputs are skipped).  Each step's
 * force evaluation is recomputed from its checkpoint and differentiated
 * with its neighbour lists held fixed; h enters through the implicit
 * function theorem on rho(h) h^D = m N_ngb / A_D.  Supported: SSPH with
 * iterative h, artificial viscosity with or without the Balsara switch,
 * periodic or open boundaries, FP64 or mixed precision (the adjoint's pair
 * arithmetic then is FP32 with FP64 accumulation); no time-dependent
 * viscosity, conductivity or gravity (SPHGPU_E_INVALID otherwise).  The
 * context is left at the final state of the rollout.
 */
int sphgpu_rollout(void *ctx, int steps, double dt);
int sphgpu_rollout_adjoint(void *ctx, const double *dl_dx, const double *dl_dv, const double *dl_du,
                           double *dx0, double *dv0, double *du0);

/*
 * TimeStep::calculation over caller-owned device arrays, allocating nothing.
 * No counterpart in the reference: this is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  It
 * runs the kernels of sphgpu_timestep on `stream` (NULL: the null stream)
 * and does not synchronise:
 *   *dt = min(cfl_sound * *h_per_v_sig, cfl_force * min_i sqrt(h_i / |a_i|))
 * over the particles with |a_i| > 0 (src/timestep.cpp:18-38).  sml and the
 * acceleration components (acc_y for dim >= 2, acc_z for dim 3) are n
 * doubles each, in any particle order; h_per_v_sig is one device double;
 * dt receives one device double.  The block partials live in a 1024-double
 * array that the library holds per GPU for the life of the process, so calls
 * that overlap on one GPU (different streams or host threads) race.
 * n must be in [1, SPHGPU_NOALLOC_MAX_N]; SPHGPU_E_INVALID otherwise.
 */
#define SPHGPU_NOALLOC_MAX_N ((int64_t)1 << 30)
int sphgpu_timestep_noalloc(int dim, int64_t n, const double *sml, const double *acc_x, const double *acc_y,
                            const double *acc_z, const double *h_per_v_sig, double cfl_sound, double cfl_force,
                            double *dt, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SPHGPU_H */
