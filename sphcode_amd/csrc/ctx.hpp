// Device context: particle state (structure of arrays, FP64 master copy in
// Morton order), the octree, neighbour lists and counters, plus the host
// launchers of every stage.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <string>

#include "../../include/sphgpu.h"
#include "common.hpp"

namespace sphgpu {

// Per-particle FP64 fields.  Vectors occupy three consecutive slots
// (axis k at base + k); only the first D are allocated.
enum Field : int {
    F_X = 0, F_V = 3, F_VP = 6, F_A = 9,
    F_MASS = 12, F_DENS, F_PRES, F_ENE, F_ENEP, F_DENE, F_SML, F_SOUND,
    F_BAL, F_ALPHA, F_GRADH, F_PHI,
    F_GD = 24, F_GP = 27, F_GV = 30,   // GSPH gradients; grad v_k along l at F_GV + 3k + l
    F_HINV = 39,                        // 1/h, written with h by the pre-interaction
    NF = 40
};

// Kernel-argument view of the state.
struct State {
    double *f[NF];
    float *g[NF];  // FP32 mirrors read by the mixed-precision pair kernels
                   // (positions excepted; allocated in mixed mode only)
    int *id;       // caller index of the particle in each sorted slot
    int *nb;       // neighbour count (SPHParticle::neighbor)
};

// FP64 master field, or its FP32 mirror, as the pair arithmetic type.
template <class T> __device__ __forceinline__ T fld(const State &s, int f, int i)
{
    if constexpr(sizeof(T) == 8) return s.f[f][i];
    else return s.g[f][i];
}

// Per-particle sums.  FP32 pair terms are compensated by accumulating them
// in FP64: every FP32 term is exact in FP64, so the sum's error does not
// grow with FP32 round-off times the number of terms.  On CDNA4 this is two
// VALU operations per term (convert, add) against four for Kahan's FP32
// compensation, at FP64 (rather than doubled-FP32) accuracy.
template <class T> struct Acc {
    double s = 0.0;
    __device__ __forceinline__ void add(T x) { s += (double)x; }
    // s += x * c with the FP64 factor x kept: one FP64 fused multiply-add
    __device__ __forceinline__ void add_prod(double x, T c) { s = fma(x, (double)c, s); }
    __device__ __forceinline__ double val() const { return s; }
};

// Octree in depth-first pre-order; node 0 is the root.
struct TreeDev {
    double4 *geo;    // centre (x,y,z), edge
    double4 *mom;    // centre of mass (x,y,z), mass
    double *ks;      // largest h in the subtree (BHNode::kernel_size)
    int4 *rng;       // first particle, end particle, skip (pre-order successor
                     // after the subtree), level | leaf << 8
    int *num_nodes;  // device scalar
};

// Gravity walk records, rebuilt before every gravity pass so that one
// scalar load brings all a wave needs at a node or at a leaf particle.
struct __align__(64) GravNode {
    double cx, cy, cz;   // centre of mass (zero at the root, as BHTree keeps it)
    double gm;           // G * mass
    double edge2;        // edge^2, for the opening test edge^2 > theta^2 d^2
    int begin, end;      // particle range
    int skip;            // pre-order successor after the subtree
    int leaf;
    float gmf;           // G * mass rounded to FP32 (mixed mode)
    int pad;
};
static_assert(sizeof(GravNode) == 64, "GravNode is one 64-byte record");
// An opened leaf recorded by the split gravity walk: its particle range and
// the wave's lanes that opened it.
struct __align__(16) LeafHit {
    int b, e;
    uint64_t mask;
};
constexpr int LEAFHIT_CAP = 1024;   // recorded leaves per wave before it falls back
constexpr int LEAFHIT_CHUNK = 256;  // records the pair pass takes per round
template <class T> struct GravPart;
template <> struct __align__(32) GravPart<float> {  // mixed: FP32 mirrors of m and 2/h
    double x[3];
    float m, einv;
};
template <> struct __align__(16) GravPart<double> {
    double x[3];
    double m, einv, pad;
};

// Neighbour records of the force pass, packed once per pass so that a pair
// reads one FP64 (x, y, z, h) record for the cut and one record of the pair
// fields (FP32 mirrors in mixed mode) instead of one load per field.
template <class T> struct __align__(16) ForceRec {
    T hinv, m, rho, p, c, alpha, bal, u, gradh;  // the kernels need 1/h only
    T v[3];
};
static_assert(sizeof(ForceRec<float>) == 48 && sizeof(ForceRec<double>) == 96, "ForceRec size");

// GSPH second order: the MUSCL gradients of a particle (density,
// pressure, velocity component c along l at gv[3c + l]) in one record.
template <class T> struct __align__(16) GradRec {
    T gd[3], gp[3], gv[9];
    T pad;
};
static_assert(sizeof(GradRec<float>) == 64 && sizeof(GradRec<double>) == 128, "GradRec size");

// Neighbour records of the pre-interaction's final sweep: FP64 position,
// and m, u, c, v in the pair type.
template <class T> struct PreRec;
template <> struct __align__(16) PreRec<float> {
    double x[3];
    float m, u, c, v[3];
};
template <> struct __align__(16) PreRec<double> {
    double x[3];
    double m, u, c, v[3];
};
static_assert(sizeof(PreRec<float>) == 48 && sizeof(PreRec<double>) == 80, "PreRec size");

// Register budgets, as the minimum number of waves per SIMD the compiler
// must fit (amdgpu_waves_per_eu), for kernels that wait on dependent
// gathers and where more resident waves beat a few spilled registers
// (measured; the pre-interaction and force kernels run best at the
// compiler's own choice).  FP64 gravity keeps the compiler's choice.
constexpr int OCC_SEARCH = 6;
template <class T> struct GravityOcc {
    static constexpr int value = sizeof(T) == 4 ? 8 : 1;
};

// Neighbour-list addressing.  The lists of 64 consecutive particles are
// interleaved in chunks of LIST_C entries: entry k of particle i lives at
//   ((i / 64) * (cap / LIST_C) + k / LIST_C) * 64 * LIST_C + (i % 64) * LIST_C + k % LIST_C.
// A search lane appending to its list fills whole LIST_C * 4-byte sectors,
// and a wave reading entry k of its 64 lists touches 64 * LIST_C * 4 bytes
// that serve the next LIST_C - 1 entries too.  cap is a multiple of LIST_C.
constexpr int LIST_C = 8;
__device__ __host__ __forceinline__ size_t list_at(int i, int k, int cap)
{
    return ((size_t)(i / WAVE) * (size_t)(cap / LIST_C) + (size_t)(k / LIST_C)) * (WAVE * LIST_C) +
           (size_t)(i % WAVE) * LIST_C + (size_t)(k % LIST_C);
}

// list_at(i, k, cap) == list_base(i, cap) + list_off(k) for 0 <= k < cap.
__device__ __host__ __forceinline__ size_t list_base(int i, int cap)
{
    return (size_t)(i / WAVE) * (size_t)(cap / LIST_C) * (WAVE * LIST_C) + (size_t)(i % WAVE) * LIST_C;
}
__device__ __host__ __forceinline__ unsigned list_off(int k)
{
    return ((unsigned)k / LIST_C) * (WAVE * LIST_C) + (unsigned)k % LIST_C;
}

// Slot of entry kk of a list with `front` entries at the start and the rest
// at the end (see Ctx::list).
__device__ __forceinline__ int list_slot(int kk, int front, int cap) { return kk < front ? kk : cap - 1 - (kk - front); }

// Device-side counters (int64), reduced per wave before the atomic.
enum Counter : int {
    C_GATHER_PAIRS = 0, C_FORCE_PAIRS, C_GRAV_PP, C_GRAV_PN, C_NEWTON_FAIL, C_NEWTON_ITERS,
    C_TRUNC, C_OVERFLOW, C_MAXLIST, C_TREE_ERR, C_GRAV_OPEN, C_SEARCH_NODES, C_SEARCH_CAND, C_GRAV_STEPS, C_GRAV_REWALK, C_GRAV_SLOTS,
    C_TILE_OVER, C_GRAV_MAXREC, NCOUNTERS
};

// Device scalars.
enum Scalar : int {
    S_DT = 0, S_HVS, S_ROOT_CX, S_ROOT_CY, S_ROOT_CZ, S_ROOT_EDGE, NSCALARS
};
constexpr size_t SCAL_SPAN = 256;  // bytes of the scalars' part of their allocation
static_assert(NSCALARS * sizeof(double) <= SCAL_SPAN, "scalars fit their span");

// The counters are STAT_SLOTS rows of STAT_ROW entries (256 bytes each).
// Row 0 is the counter array proper: the rare events the host checks after
// every stage (truncations, overflows, tree errors) go there.  The
// statistics that every wave adds to (pairs, iterations, node visits) go
// to the row of the wave's task modulo STAT_SLOTS, so that the waves'
// atomic updates spread over many cache lines instead of queueing on one
// (measured: the Evrard step 4 % faster); sphgpu_get_stats sums the rows
// (maxima: takes the largest).
constexpr int STAT_SLOTS = 64;
constexpr int STAT_ROW = 32;
static_assert(NCOUNTERS <= STAT_ROW, "a counter row holds every counter");
constexpr int NCOUNTER_ENTRIES = STAT_SLOTS * STAT_ROW;

__device__ __forceinline__ void stat_add(int64_t *counters, int task, int k, int64_t v)
{
    atomicAdd((unsigned long long *)&counters[(task & (STAT_SLOTS - 1)) * STAT_ROW + k], (unsigned long long)v);
}

__device__ __forceinline__ void stat_max(int64_t *counters, int task, int k, int64_t v)
{
    atomicMax((unsigned long long *)&counters[(task & (STAT_SLOTS - 1)) * STAT_ROW + k], (unsigned long long)v);
}


constexpr int TIMER_SETS = 32;
constexpr int TIMER_EVENTS = 20;

struct Ctx {
    sphgpu_params p{};
    int D = 3;
    hipStream_t stream = nullptr;
    // second stream (default priority) on which sphgpu_integrate runs the
    // gravity pass alongside the force pass
    hipStream_t gstream = nullptr;
    hipEvent_t ev_pre = nullptr, ev_grav = nullptr;
    hipEvent_t *ev_g = nullptr;  // the gravity stream's timing events of the current timer set
    // sphgpu_integrate's check of the combined lists, read without stalling
    // the stream: the counters are copied after the pre-interaction (event
    // ev_flags) and read once the force pass is queued
    hipEvent_t ev_flags = nullptr;
    bool speculate = false;      // sphgpu_integrate: assume symmetric lists, check after
    bool reorder_light = false;  // sphgpu_integrate's tree build: skip fields the step rewrites
    bool radius_ready = false;   // that build set the search radii and the cells' largest radius
    bool stats_on = false;       // count the statistics (sphgpu_enable_stats)
    bool flags_pending = false;  // a copy behind ev_flags awaits its check
    int64_t *counters_spec = nullptr;  // pinned host copy for that check
    // called by the force pass between its record packing and its pair
    // kernel (sphgpu_integrate starts the gravity pair pass there, so that
    // the packing is not left waiting behind it); its status is kept here
    int (*after_pack)(Ctx &) = nullptr;
    int after_pack_rc = 0;
    int device = 0;
    int num_cu = 256;         // compute units of the device
    int *work = nullptr;      // per-XCD task counters of persistent kernels

    int64_t n = 0;        // particles
    int64_t cap_n = 0;    // allocated particles
    int cap = 0;          // per-particle list capacity
    int64_t node_cap = 0; // allocated nodes

    State s{};       // current buffers
    State alt{};     // reorder targets

    // tree build
    uint64_t *keys = nullptr, *keys_sorted = nullptr;
    int *idx = nullptr, *perm = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    uint8_t *delta = nullptr;
    int *nn = nullptr, *noff = nullptr;
    void *scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    TreeDev t{};
    int *lvl_nodes = nullptr;  // node ids grouped by tree level
    int *lvl = nullptr;        // per-level node counts, offsets, fill (device)
    int *lvl_host = nullptr;   // per-level counts and offsets (pinned host)
    double4 *msum = nullptr;   // raw mass moments (sum m x, sum m)
    GravNode *gnode = nullptr; // gravity node records (gravity on)
    double *gacc = nullptr;    // gravity's acceleration, D x cap_n (gravity on)
    LeafHit *leafhit = nullptr; // split walk: LEAFHIT_CAP records per wave
    int *leafcnt = nullptr;     // split walk: records per wave
    double4 *posrec = nullptr; // force pass: x, y, z, h per particle
    void *pair_rec = nullptr;  // records of the running pair pass: PreRec<T>
                               // (pre-interaction) or ForceRec<T> (force)
    void *grad_rec = nullptr;  // GSPH second order: GradRec<T> per particle
    void *gpart = nullptr;     // GravPart<T> per particle (gravity on)
    // search groups: particle ranges of at most 64 particles, each inside
    // the deepest cell holding at most 64 particles
    int2 *groups = nullptr;
    int *num_groups = nullptr;   // device scalar (an entry of lvl)
    int64_t group_cap = 0;

    // candidate tiles of the last neighbour search, per search group g
    // (groups[g], at most 64 particles): tile_nc[g] candidates, stored in
    // the group's slab of tile_wcap slots starting at g * tile_wcap.  With
    // T = ceil(nc / 64) tiles, staging-order candidate c (the order of the
    // reference's one-leaf-at-a-time walk) is lane k = c / T of tile
    // t = c % T: tile t's 64 candidate indices sit at slab + 64 t (-1 past
    // the end), and lane l of the group has the bit k of tile_mf[slab + 64 t
    // + l] set when candidate (t, k) is in its front set (the gather
    // candidates, r < R_l) and of tile_ma[...] when it is any of its hits
    // (front or other symmetric candidates).  Consecutive candidates land in different
    // tiles, so each tile samples the whole neighbourhood of the group and
    // the group's lanes hold similar numbers of hits in every tile.
    int *tile_cand = nullptr;
    uint64_t *tile_mf = nullptr, *tile_ma = nullptr;
    int *tile_nc = nullptr;
    int tile_wcap = 0;             // slots per group (multiple of 64)
    int64_t tile_groups_cap = 0;   // groups the slabs are allocated for
    int *ws_j = nullptr;           // search waves' staging of kept candidates:
    double4 *ws_p = nullptr;       // index, and position with radius squared
    int64_t ws_waves = 0;          // search waves the staging is allocated for
    int ngroups = 0;               // groups of the last tree (host copy)

    // neighbour lists (addressing: list_at), expanded from the tiles when
    // the adjoint or the set record needs them (need_lists).  The first
    // list_cnt[i] entries (k = 0, 1, ...) are the gather candidates
    // (r < h_guess * ratio) with |r_ij| in list_r / list_rf; the combined
    // search appends the remaining symmetric candidates from the other end
    // (k = cap-1, cap-2, ..., list_cnt2[i] of them).
    int *list = nullptr;
    double *list_r = nullptr;  // |r_ij| of gather candidates (FP64 mode)
    float *list_rf = nullptr;  // |r_ij| of gather candidates (mixed mode)
    int *list_cnt = nullptr;
    int *list_cnt2 = nullptr;
    bool need_lists = false;       // expand every search's tiles into list
    bool lists_symmetric = false;  // the tiles hold every force pair
    int64_t trunc_seen = 0;        // truncation counter at the last check
    int *rec_gather = nullptr, *rec_gather_cnt = nullptr;  // recorded sets (debug)
    int *rec_force = nullptr, *rec_force_cnt = nullptr;
    bool record = false;
    bool have_gather = false, have_force = false;

    // device scalars, and the counters SCAL_SPAN bytes further on in the
    // same allocation (one copy brings both to the host; the counters, which
    // kernels update with atomics, stay off the cache lines of the scalars,
    // which every kernel reads), and their pinned host copies laid out alike
    double *scal = nullptr;
    double *partial = nullptr;  // block partials of reductions
    int64_t *counters = nullptr;       // (char *)scal + SCAL_SPAN
    int64_t *counters_host = nullptr;  // (char *)scal_host + SCAL_SPAN
    double *scal_host = nullptr;

    // host mirror of simulation scalars
    double time = 0.0, dt = 0.0, h_per_v_sig = 0.0;
    bool uploaded = false, first_pre = true, tree_ready = false;
    bool gpart_fresh = false;  // the last pre-interaction wrote gpart for this tree

    // differentiable rollout: per-step checkpoints (device order, with the
    // id permutation of that moment) and adjoint buffers
    struct Rollout {
        int steps = 0;          // K of the stored rollout (0: none)
        double dt = 0.0;
        double t0 = 0.0;
        int64_t n = 0;
        double *ck = nullptr;   // (K + 2) slots x NCK fields x n
        int *ck_id = nullptr;   // (K + 2) slots x n
        double *adj = nullptr;  // adjoint of the state, caller order
        double *work = nullptr; // per-stage adjoint arrays, device order
        int slots = 0;          // allocated slots
    } ro;

    // stage timers
    bool timers_on = false;
    bool kernel_marks = false;  // sphgpu_integrate with timers on: mark kernel spans
    // sphgpu_integrate's timers: each timed step records into one of
    // TIMER_SETS sets of events (ev[0..7] stage boundaries, ev[8..13]
    // kernel spans, then the gravity stream's four), and the spans are
    // summed into timer_ms only when the timers are read or the sets run
    // out, so that no step waits on the host reading the previous one's
    hipEvent_t tev[TIMER_SETS][TIMER_EVENTS] = {};  // created when timers are first enabled
    bool tset_grav[TIMER_SETS] = {};
    int tset_cur = 0, tset_pending = 0;  // pending sets end at tset_cur
    hipEvent_t *ev = nullptr;            // the current set
    double timer_ms[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int64_t steps = 0;
    int64_t last_nodes = 0, last_levels = 0;
};

// Error state (per thread).
int set_error(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define SPHGPU_HIP(call)                                              \
    do {                                                              \
        hipError_t e_ = (call);                                       \
        if(e_ != hipSuccess) return ::sphgpu::hip_fail(e_, #call);    \
    } while(0)

inline int grid_for(int64_t n, int block = BLOCK) { return (int)((n + block - 1) / block); }

BoxDev make_box(const Ctx &c);

// ------------------------------------------------------------- launchers
// tree.hip
int tree_alloc(Ctx &c);
int tree_build(Ctx &c);          // BHTree::make
int tree_set_kernel(Ctx &c);     // BHTree::set_kernel
// search.hip: candidate lists from the tree walk
int search_gather(Ctx &c, bool initial);  // h from the density guess
int search_symmetric(Ctx &c);
int search_guess(Ctx &c);        // sml = h_guess * ratio
int search_combined(Ctx &c, bool work_ready = false, int blocks_per_cu = 8);  // gather and symmetric candidates in one walk
int search_alloc(Ctx &c);        // tile slabs for c.ngroups groups (after a tree build)
int search_expand(Ctx &c, int *list, int *cnt, int *cnt2, bool front_only);  // tiles -> list_at lists
// pre.hip
int pre_initial(Ctx &c);         // initial_smoothing density
int pre_compute(Ctx &c);         // Newton h, density, pressure, grad-h, AV switches / MUSCL gradients
// force.hip
int force_compute(Ctx &c);
// gravity.hip
// The gravity pass in three parts, each on stream st: the walk (monopoles
// into phi and gacc, and a record of the opened leaves; needs positions and
// the tree only), the recorded leaves' particle pairs (need h; added to
// phi and gacc), and the combine acc += gacc (after the force pass has
// assigned acc).  gravity_compute runs all three on c.stream;
// sphgpu_integrate overlaps the first two with the hydrodynamic stages.
int gravity_compute(Ctx &c);
int gravity_walk(Ctx &c, hipStream_t st);
int gravity_leaves(Ctx &c, hipStream_t st);
int gravity_combine(Ctx &c, hipStream_t st);
// Zero the NUM_XCD task counters at c.work + NUM_XCD * slot before a
// persistent launch (slot 0: the neighbour search).
void reset_work(Ctx &c, int slot);
constexpr int WORK_SLOTS = 1;
// one wave per search group: grid of ceil(ngroups / 4) blocks of 4 waves,
// block b serving groups 4 xcd_block() .. + 3
inline int group_grid(const Ctx &c) { return std::max(1, (c.ngroups + BLOCK / WAVE - 1) / (BLOCK / WAVE)); }
// step.hip
int step_init_fields(Ctx &c);
int step_timestep(Ctx &c);       // dt into scal[S_DT]
int step_predict(Ctx &c);
int step_correct(Ctx &c, bool add_gravity = false);  // add_gravity: acc += the gravity pass's sums first
int step_energy(Ctx &c, double *out);
// FP32 mirrors of the listed fields (mixed precision only; no-op in FP64)
int mirror_fields(Ctx &c, const int *fields, int nfields);
int reduce_min_partials(Ctx &c, int nblocks, int slot);
// adjoint.hip: vector-Jacobian product of one force evaluation
// (pre-interaction + fluid force at the current state and lists) and the
// kick-drift around it; kick = false for the evaluation of Solver::initialize
int adjoint_step(Ctx &c, bool kick);
int adjoint_alloc(Ctx &c, int steps);
void adjoint_free(Ctx &c);
int adjoint_save(Ctx &c, int slot);
int adjoint_restore(Ctx &c, int slot);
int adjoint_seed(Ctx &c, const double *dx, const double *dv, const double *du);
int adjoint_fetch(Ctx &c, double *dx, double *dv, double *du);

}  // namespace sphgpu
