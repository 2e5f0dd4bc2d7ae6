// The extern "C" entry points of include/sphgpu.h.  Each call validates its
// arguments, enqueues the stage's kernels on the context stream and reports
// failures as negative status codes with a message (the reference throws
// SPHException, include/exception.hpp:9-14).
#include "ctx.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace sphgpu {

static thread_local std::string g_error;

// Events that only order the two streams or time spans need the data before
// them visible to this device, not to the host.
#ifndef EV_SCOPE
#define EV_SCOPE hipEventReleaseToDevice
#endif

int set_error(int code, const std::string &msg)
{
    g_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return set_error(SPHGPU_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

__global__ void set_scalar_kernel(double *scal, int slot, double v) { scal[slot] = v; }

__global__ void fill_int_kernel(int *p, int64_t n, int v)
{
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if(i < n) p[i] = v;
}

void reset_work(Ctx &c, int slot)
{
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(BLOCK), 0, c.stream, c.work + NUM_XCD * slot, (int64_t)NUM_XCD, 0);
}

__global__ void iota_kernel(int *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if(i < n) p[i] = (int)i;
}

__global__ void fill_double_kernel(double *p, int64_t n, double v)
{
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if(i < n) p[i] = v;
}

__global__ void fill_i64_kernel(int64_t *p, int n)
{
    for(int k = threadIdx.x; k < n; k += blockDim.x) p[k] = 0;
}

static bool field_used(const Ctx &c, int f)
{
    const int D = c.D;
    if(f < F_MASS) return (f % 3) < D;
    if(f >= F_GD && f < F_GD + 3) return c.p.sph_type == SPHGPU_GSPH && f - F_GD < D;
    if(f >= F_GP && f < F_GP + 3) return c.p.sph_type == SPHGPU_GSPH && f - F_GP < D;
    if(f >= F_GV && f < F_GV + 9) return c.p.sph_type == SPHGPU_GSPH && (f - F_GV) / 3 < D && (f - F_GV) % 3 < D;
    return true;
}

static void free_buffers(Ctx &c)
{
    auto fr = [](auto *&p) {
        if(p) (void)hipFree((void *)p);
        p = nullptr;
    };
    for(int f = 0; f < NF; ++f) {
        fr(c.s.f[f]);
        fr(c.alt.f[f]);
        fr(c.s.g[f]);
        c.alt.g[f] = nullptr;
    }
    fr(c.list_rf);
    fr(c.s.id); fr(c.alt.id); fr(c.s.nb); fr(c.alt.nb);
    fr(c.keys); fr(c.keys_sorted); fr(c.idx); fr(c.perm); fr(c.sort_tmp); fr(c.delta); fr(c.nn); fr(c.noff);
    fr(c.scan_tmp); fr(c.t.geo); fr(c.t.mom); fr(c.t.ks); fr(c.t.rng); fr(c.t.num_nodes);
    fr(c.lvl_nodes); fr(c.msum); fr(c.groups); fr(c.lvl);
    c.num_groups = nullptr;  // inside lvl
    fr(c.gnode); fr(c.gpart); fr(c.gacc); fr(c.leafhit); fr(c.leafcnt); fr(c.posrec); fr(c.pair_rec); fr(c.grad_rec);
    if(c.lvl_host) (void)hipHostFree(c.lvl_host);
    c.lvl_host = nullptr;
    fr(c.list); fr(c.list_r); fr(c.list_cnt); fr(c.list_cnt2);
    fr(c.tile_cand); fr(c.tile_mf); fr(c.tile_ma); fr(c.tile_nc); fr(c.ws_j); fr(c.ws_p);
    c.tile_wcap = 0;
    c.tile_groups_cap = 0;
    c.ws_waves = 0;
    c.ngroups = 0;
    fr(c.rec_gather); fr(c.rec_gather_cnt); fr(c.rec_force); fr(c.rec_force_cnt);
    fr(c.partial);
    c.cap_n = 0;
}

static int alloc_buffers(Ctx &c, int64_t n)
{
    free_buffers(c);
    const int64_t np = (n + WAVE - 1) / WAVE * WAVE;
    c.cap_n = np;
    const bool mixed = c.p.precision == SPHGPU_MIXED;
    for(int f = 0; f < NF; ++f) {
        if(!field_used(c, f)) continue;
        SPHGPU_HIP(hipMalloc(&c.s.f[f], np * sizeof(double)));
        SPHGPU_HIP(hipMalloc(&c.alt.f[f], np * sizeof(double)));
        if(mixed && f >= F_V) SPHGPU_HIP(hipMalloc(&c.s.g[f], np * sizeof(float)));
    }
    SPHGPU_HIP(hipMalloc(&c.posrec, np * sizeof(double4)));
    // one buffer for the records of either pair pass (ForceRec is the larger)
    static_assert(sizeof(ForceRec<float>) >= sizeof(PreRec<float>) && sizeof(ForceRec<double>) >= sizeof(PreRec<double>),
                  "pair_rec holds either record");
    SPHGPU_HIP(hipMalloc(&c.pair_rec, np * (mixed ? sizeof(ForceRec<float>) : sizeof(ForceRec<double>))));
    if(c.p.sph_type == SPHGPU_GSPH && c.p.gsph_is_2nd_order)
        SPHGPU_HIP(hipMalloc(&c.grad_rec, np * (mixed ? sizeof(GradRec<float>) : sizeof(GradRec<double>))));
    SPHGPU_HIP(hipMalloc(&c.s.id, np * sizeof(int)));
    SPHGPU_HIP(hipMalloc(&c.alt.id, np * sizeof(int)));
    SPHGPU_HIP(hipMalloc(&c.s.nb, np * sizeof(int)));
    SPHGPU_HIP(hipMalloc(&c.alt.nb, np * sizeof(int)));
    // per-particle lists are expanded from the candidate tiles only for the
    // adjoint (allocated then, ensure_lists)
    const size_t lsz = (size_t)np * (size_t)c.cap;
    SPHGPU_HIP(hipMalloc(&c.list_cnt, np * sizeof(int)));
    SPHGPU_HIP(hipMalloc(&c.list_cnt2, np * sizeof(int)));
    // block partials: per 256 particles, or per block of four search groups
    // (up to one group per particle)
    const int64_t npart = std::max<int64_t>(grid_for(np), 1024) * 8 + np / 4 + 64;
    SPHGPU_HIP(hipMalloc(&c.partial, npart * sizeof(double)));
    int rc = tree_alloc(c);
    if(rc) return rc;
    if(c.record) {
        SPHGPU_HIP(hipMalloc(&c.rec_gather, lsz * sizeof(int)));
        SPHGPU_HIP(hipMalloc(&c.rec_gather_cnt, np * sizeof(int)));
        SPHGPU_HIP(hipMalloc(&c.rec_force, lsz * sizeof(int)));
        SPHGPU_HIP(hipMalloc(&c.rec_force_cnt, np * sizeof(int)));
    }
    return SPHGPU_OK;
}

// Reads the error counters; a list overflow or node-pool exhaustion since
// the last check becomes an error status.
// Neighbour lists in list_at layout (the adjoint's input), allocated on
// first use.
static int ensure_lists(Ctx &c)
{
    if(!c.list) SPHGPU_HIP(hipMalloc(&c.list, (size_t)c.cap_n * (size_t)c.cap * sizeof(int)));
    return SPHGPU_OK;
}

// Reads the counters (and, with the scalars, the scalars as well: they
// share one allocation, so one copy brings both) and reports the errors
// they record.
static int check_flags(Ctx &c, const char *stage, bool with_scalars = false)
{
    if(with_scalars)
        SPHGPU_HIP(hipMemcpyAsync(c.scal_host, c.scal, SCAL_SPAN + NCOUNTERS * sizeof(int64_t), hipMemcpyDeviceToHost,
                                  c.stream));
    else
        SPHGPU_HIP(hipMemcpyAsync(c.counters_host, c.counters, NCOUNTERS * sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    if(c.counters_host[C_TREE_ERR] > 0)
        return set_error(SPHGPU_E_TREE, std::string(stage) + ": There is no free node.");
    if(c.counters_host[C_OVERFLOW] > 0)
        return set_error(SPHGPU_E_LIST_OVERFLOW, std::string(stage) + ": neighbour list capacity (" + std::to_string(c.cap) +
                                                     ") exceeded; raise neighbor_capacity");
    if(c.counters_host[C_TILE_OVER] > 0)
        return set_error(SPHGPU_E_LIST_OVERFLOW, std::string(stage) + ": candidates of a search group exceed " +
                                                     std::to_string(c.tile_wcap) + "; raise neighbor_capacity");
    return SPHGPU_OK;
}

static int stage_tree(Ctx &c)
{
    int rc = tree_build(c);
    if(rc) return rc;
    c.tree_ready = true;
    c.gpart_fresh = false;
    c.have_gather = c.have_force = false;
    c.lists_symmetric = false;  // lists refer to the previous order
    return SPHGPU_OK;
}

// sphgpu_integrate's gravity walk on the second stream, ordered after the
// work queued so far on the main stream.
static int launch_gravity_walk(Ctx &c)
{
    SPHGPU_HIP(hipEventRecord(c.ev_pre, c.stream));
    SPHGPU_HIP(hipStreamWaitEvent(c.gstream, c.ev_pre, 0));
    if(c.timers_on) (void)hipEventRecord(c.ev_g[0], c.gstream);
    int rc = gravity_walk(c, c.gstream);
    if(c.timers_on) (void)hipEventRecord(c.ev_g[1], c.gstream);
    SPHGPU_HIP(hipEventRecord(c.ev_grav, c.gstream));
    return rc;
}

// sphgpu_integrate's gravity pair pass on the second stream, after the
// main stream's work so far (the pre-interaction's h and the force pass's
// packing) and the walk.
static int launch_gravity_leaves(Ctx &c)
{
    SPHGPU_HIP(hipEventRecord(c.ev_pre, c.stream));
    SPHGPU_HIP(hipStreamWaitEvent(c.gstream, c.ev_pre, 0));
    if(c.timers_on) (void)hipEventRecord(c.ev_g[2], c.gstream);
    int rc = gravity_leaves(c, c.gstream);
    if(c.timers_on) (void)hipEventRecord(c.ev_g[3], c.gstream);
    SPHGPU_HIP(hipEventRecord(c.ev_grav, c.gstream));
    return rc;
}

// The pre-interaction's candidate search.  After the first call, one walk
// finds both the gather candidates and the symmetric pairs the force stage
// needs (search_combined); the first call keeps the reference's initial
// smoothing and separate searches.  With `walk` the gravity walk is started
// on the second stream just before the search (after the search radii, the
// cells' largest radius and the search's task counters are set, so that no
// small kernel of the main stream waits for CUs behind the walk).
// Kernel-span marks of sphgpu_integrate's timers (ev[8 + k]).
static void kmark(Ctx &c, int k)
{
    if(c.kernel_marks) (void)hipEventRecord(c.ev[8 + k], c.stream);
}

static int stage_pre(Ctx &c, bool walk = false)
{
    int rc;
    const bool combined = !c.first_pre;
    if(c.first_pre) {
        // PreInteraction::initial_smoothing (src/pre_interaction.cpp:41-44, 171-215)
        if((rc = search_gather(c, true))) return rc;
        if((rc = pre_initial(c))) return rc;
        c.first_pre = false;
        kmark(c, 0);
        if((rc = search_gather(c, false))) return rc;
        kmark(c, 1);
    } else {
        // the radii h_guess * ratio and the largest radius per cell (set by
        // sphgpu_integrate's tree build)
        if(!c.radius_ready && ((rc = search_guess(c)) || (rc = tree_set_kernel(c)))) return rc;
        c.radius_ready = false;
        reset_work(c, 0);  // the search's task counters
        // The gravity walk (second stream) runs beside the search.  The
        // search then takes two blocks of four waves per CU, leaving the
        // other wave slots to the walk: the search waits on dependent loads
        // and gains little from more waves, while the walk is bound by
        // instruction issue (measured per mixed Evrard step of 2^20
        // particles: 7.04 ms, against 7.16 ms with three blocks per CU and
        // 7.35 ms with four).  Without gravity the best share depends on
        // the dimension: the 2D KHI search (393,160 particles) takes
        // 0.56 ms with two blocks per CU against 0.82 ms with eight (fewer
        // waves keep their staging and tiles in L2), while the 3D Sedov
        // step (2^24 particles) takes 65.8 ms with eight against 68.4 ms
        // with four and 87.4 ms with two (its tree walks need the waves to
        // hide their dependent loads).
        const int bpc = (walk || c.D < 3) ? 2 : 8;
        if(walk && (rc = launch_gravity_walk(c))) return rc;
        walk = false;
        kmark(c, 0);
        if((rc = search_combined(c, true, bpc))) return rc;
        kmark(c, 1);
    }
    if(c.record) {
        if((rc = search_expand(c, c.rec_gather, c.rec_gather_cnt, nullptr, true))) return rc;
        c.have_gather = true;
    }
    if(c.need_lists && ((rc = ensure_lists(c)) || (rc = search_expand(c, c.list, c.list_cnt, c.list_cnt2, false))))
        return rc;
    if(walk && (rc = launch_gravity_walk(c))) return rc;
    kmark(c, 2);
    if((rc = pre_compute(c))) return rc;
    kmark(c, 3);
    c.lists_symmetric = false;
    if(combined) {
        // the combined lists hold every force pair unless some h outgrew its
        // search radius (src/pre_interaction.cpp:70 truncates the gather set)
        if(c.speculate && !c.need_lists && !c.record) {
            // sphgpu_integrate: the force pass starts on the combined lists
            // at once and stage_force checks the counters once it is queued
            // (a truncation, which needs a new search, is rare)
            SPHGPU_HIP(hipMemcpyAsync(c.counters_spec, c.counters, NCOUNTERS * sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
            SPHGPU_HIP(hipEventRecord(c.ev_flags, c.stream));
            c.flags_pending = true;
            c.lists_symmetric = true;
            return SPHGPU_OK;
        }
        SPHGPU_HIP(hipMemcpyAsync(c.counters_host, c.counters, NCOUNTERS * sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
        SPHGPU_HIP(hipStreamSynchronize(c.stream));
        const int64_t tr = c.counters_host[C_TRUNC];
        c.lists_symmetric = tr == c.trunc_seen && c.counters_host[C_OVERFLOW] == 0;
        c.trunc_seen = tr;
    }
    if(c.lists_symmetric) return SPHGPU_OK;
    return tree_set_kernel(c);  // BHTree::set_kernel (src/pre_interaction.cpp:166-168)
}

static int stage_force(Ctx &c)
{
    int rc;
    if(!c.lists_symmetric) {
        if((rc = search_symmetric(c))) return rc;
        if(c.need_lists && ((rc = ensure_lists(c)) || (rc = search_expand(c, c.list, c.list_cnt, c.list_cnt2, false))))
            return rc;
    }
    c.lists_symmetric = false;
    kmark(c, 4);
    if((rc = force_compute(c))) return rc;
    kmark(c, 5);
    if(c.flags_pending) {
        // the speculative force pass is queued: check the combined lists now
        c.flags_pending = false;
        SPHGPU_HIP(hipEventSynchronize(c.ev_flags));
        const int64_t tr = c.counters_spec[C_TRUNC];
        const bool ok = tr == c.trunc_seen && c.counters_spec[C_OVERFLOW] == 0;
        c.trunc_seen = tr;
        if(!ok) {
            // some h outgrew its search radius: the cells' largest h, a
            // symmetric search, and the force pass again (it assigns acc and
            // du/dt in full; the gravity pair pass started by the first one
            // is not started again)
            int (*hook)(Ctx &) = c.after_pack;
            c.after_pack = nullptr;
            if((rc = tree_set_kernel(c)) || (rc = search_symmetric(c)) || (rc = force_compute(c))) {
                c.after_pack = hook;
                return rc;
            }
            c.after_pack = hook;
        }
    }
    if(c.record) c.have_force = true;
    return SPHGPU_OK;
}

}  // namespace sphgpu

using namespace sphgpu;

static Ctx *as_ctx(void *p) { return static_cast<Ctx *>(p); }

#define SPHGPU_CTX(ptr)                                                        \
    if(!(ptr)) return set_error(SPHGPU_E_INVALID, "null context");           \
    Ctx &c = *as_ctx(ptr);                                                     \
    if(hipSetDevice(c.device) != hipSuccess) return set_error(SPHGPU_E_HIP, "hipSetDevice failed")

#define SPHGPU_NEED_DATA()                                                     \
    if(!c.uploaded) return set_error(SPHGPU_E_STATE, "no particles uploaded")

#define SPHGPU_NEED_TREE()                                                     \
    SPHGPU_NEED_DATA();                                                        \
    if(!c.tree_ready) return set_error(SPHGPU_E_STATE, "tree not built")

extern "C" {

const char *sphgpu_last_error(void) { return g_error.c_str(); }

int sphgpu_version(void) { return 1; }

int sphgpu_params_default(sphgpu_params *p)
{
    if(!p) return set_error(SPHGPU_E_INVALID, "null params");
    std::memset(p, 0, sizeof(*p));
    p->dim = 3;
    p->sph_type = SPHGPU_SSPH;
    p->kernel = SPHGPU_CUBIC_SPLINE;
    p->precision = SPHGPU_FP64;
    p->cfl_sound = 0.3;
    p->cfl_force = 0.125;
    p->av_alpha = 1.0;
    p->av_use_balsara_switch = 1;
    p->av_use_time_dependent = 0;
    p->av_alpha_max = 2.0;
    p->av_alpha_min = 0.1;
    p->av_epsilon = 0.2;
    p->ac_alpha = 1.0;
    p->ac_is_valid = 0;
    p->tree_max_level = 20;
    p->tree_leaf_particle_num = 1;
    p->neighbor_number = 32;
    p->iterative_sml = 1;
    p->gravity_constant = 1.0;
    p->gravity_theta = 0.5;
    p->gsph_is_2nd_order = 1;
    return SPHGPU_OK;
}

int sphgpu_create(const sphgpu_params *params, void **out)
{
    if(!params || !out) return set_error(SPHGPU_E_INVALID, "null argument");
    const sphgpu_params &p = *params;
    if(p.dim < 1 || p.dim > 3) return set_error(SPHGPU_E_INVALID, "dim must be 1, 2 or 3");
    if(p.sph_type < 0 || p.sph_type > 2) return set_error(SPHGPU_E_INVALID, "Unknown SPH type");
    if(p.kernel != SPHGPU_CUBIC_SPLINE && p.kernel != SPHGPU_WENDLAND) return set_error(SPHGPU_E_INVALID, "kernel is unknown.");
    // C4Kernel asserts DIM != 1 (include/kernel/wendland_kernel.hpp:25-28):
    // its 1D normalisation is 0, so every density would be 0.
    if(p.kernel == SPHGPU_WENDLAND && p.dim == 1) return set_error(SPHGPU_E_INVALID, "Wendland C4 kernel is not defined for DIM = 1");
    if(p.precision != SPHGPU_FP64 && p.precision != SPHGPU_MIXED) return set_error(SPHGPU_E_INVALID, "unknown precision");
    if(p.tree_max_level < 1 || p.tree_max_level * p.dim > 63)
        return set_error(SPHGPU_E_INVALID, "maxTreeLevel must satisfy 1 <= maxTreeLevel and maxTreeLevel * DIM <= 63");
    if(p.tree_leaf_particle_num < 1) return set_error(SPHGPU_E_INVALID, "leafParticleNumber must be >= 1");
    if(p.neighbor_number < 1) return set_error(SPHGPU_E_INVALID, "neighborNumber must be >= 1");
    if(p.av_use_time_dependent && p.av_alpha_max < p.av_alpha_min) return set_error(SPHGPU_E_INVALID, "alphaMax < alphaMin");
    if(p.time_end < p.time_start) return set_error(SPHGPU_E_INVALID, "endTime < startTime");
    if(p.periodic_is_valid)
        for(int k = 0; k < p.dim; ++k)
            if(!(p.range_max[k] > p.range_min[k])) return set_error(SPHGPU_E_INVALID, "periodic range is empty");
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(SPHGPU_E_NO_DEVICE, "no HIP device");
    Ctx *c = new Ctx();
    c->p = p;
    c->D = p.dim;
    c->cap = p.neighbor_capacity > 0 ? p.neighbor_capacity : p.neighbor_number * 20;
    c->cap = (c->cap + LIST_C - 1) / LIST_C * LIST_C;  // whole chunks (list_at)
    c->time = p.time_start;
    c->dt = 0.0;
    int dev = 0;
    (void)hipGetDevice(&dev);
    c->device = dev;
    int cus = 0;
    if(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0) c->num_cu = cus;
    hipError_t e;
    if((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
       (e = hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking)) != hipSuccess ||
       (e = hipEventCreateWithFlags(&c->ev_pre, hipEventDisableTiming | EV_SCOPE)) != hipSuccess ||
       (e = hipEventCreateWithFlags(&c->ev_grav, hipEventDisableTiming | EV_SCOPE)) != hipSuccess ||
       (e = hipEventCreateWithFlags(&c->ev_flags, hipEventDisableTiming)) != hipSuccess ||
       (e = hipHostMalloc(&c->counters_spec, NCOUNTERS * sizeof(int64_t), hipHostMallocDefault)) != hipSuccess ||
       // scalars and counters in one allocation (device and pinned host)
       (e = hipMalloc(&c->scal, SCAL_SPAN + NCOUNTER_ENTRIES * sizeof(int64_t))) != hipSuccess ||
       (e = hipMalloc(&c->work, WORK_SLOTS * NUM_XCD * sizeof(int))) != hipSuccess ||
       (e = hipHostMalloc(&c->scal_host, SCAL_SPAN + NCOUNTER_ENTRIES * sizeof(int64_t), hipHostMallocDefault)) != hipSuccess) {
        delete c;
        return hip_fail(e, "sphgpu_create");
    }
    c->counters = reinterpret_cast<int64_t *>(reinterpret_cast<char *>(c->scal) + SCAL_SPAN);
    c->counters_host = reinterpret_cast<int64_t *>(reinterpret_cast<char *>(c->scal_host) + SCAL_SPAN);

    hipLaunchKernelGGL(fill_i64_kernel, dim3(1), dim3(BLOCK), 0, c->stream, c->counters, NCOUNTER_ENTRIES);
    hipLaunchKernelGGL(fill_double_kernel, dim3(1), dim3(BLOCK), 0, c->stream, c->scal, (int64_t)NSCALARS, 0.0);
    if(p.periodic_is_valid) {
        // BHTree::initialize (src/bhtree.cpp:19-31): root = the periodic box
        double edge = 0.0;
        for(int k = 0; k < p.dim; ++k) {
            hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, c->stream, c->scal, S_ROOT_CX + k,
                               (p.range_max[k] + p.range_min[k]) * 0.5);
            const double r = p.range_max[k] - p.range_min[k];
            if(edge < r) edge = r;
        }
        hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, c->stream, c->scal, (int)S_ROOT_EDGE, edge);
    }
    if((e = hipStreamSynchronize(c->stream)) != hipSuccess) {
        delete c;
        return hip_fail(e, "sphgpu_create");
    }
    *out = c;
    return SPHGPU_OK;
}

int sphgpu_destroy(void *ptr)
{
    if(!ptr) return SPHGPU_OK;
    Ctx *c = as_ctx(ptr);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if(c->gstream) (void)hipStreamSynchronize(c->gstream);
    adjoint_free(*c);
    free_buffers(*c);
    if(c->scal) (void)hipFree(c->scal);  // with the counters
    if(c->work) (void)hipFree(c->work);
    if(c->counters_spec) (void)hipHostFree(c->counters_spec);
    if(c->scal_host) (void)hipHostFree(c->scal_host);  // with counters_host
    for(int t = 0; t < TIMER_SETS; ++t)
        for(int k = 0; k < TIMER_EVENTS; ++k)
            if(c->tev[t][k]) (void)hipEventDestroy(c->tev[t][k]);
    if(c->stream) (void)hipStreamDestroy(c->stream);
    if(c->gstream) (void)hipStreamDestroy(c->gstream);
    for(hipEvent_t e : {c->ev_pre, c->ev_grav, c->ev_flags})
        if(e) (void)hipEventDestroy(e);
    delete c;
    return SPHGPU_OK;
}

int sphgpu_set_record_sets(void *ptr, int enable)
{
    SPHGPU_CTX(ptr);
    c.record = enable != 0;
    if(c.uploaded && c.record && !c.rec_gather) {
        const size_t lsz = (size_t)c.cap_n * (size_t)c.cap;
        SPHGPU_HIP(hipMalloc(&c.rec_gather, lsz * sizeof(int)));
        SPHGPU_HIP(hipMalloc(&c.rec_gather_cnt, c.cap_n * sizeof(int)));
        SPHGPU_HIP(hipMalloc(&c.rec_force, lsz * sizeof(int)));
        SPHGPU_HIP(hipMalloc(&c.rec_force_cnt, c.cap_n * sizeof(int)));
    }
    return SPHGPU_OK;
}

int sphgpu_upload(void *ptr, const sphgpu_particles *h)
{
    SPHGPU_CTX(ptr);
    if(!h || h->n <= 0) return set_error(SPHGPU_E_INVALID, "empty particle array");
    if(h->n > (int64_t)0x7fffffff / 2) return set_error(SPHGPU_E_INVALID, "too many particles");
    if(!h->pos || !h->mass) return set_error(SPHGPU_E_INVALID, "pos and mass are required");
    const int64_t n = h->n;
    const int D = c.D;
    if(n != c.n || !c.s.f[F_X]) {
        int rc = alloc_buffers(c, n);
        if(rc) return rc;
    }
    c.n = n;
    std::vector<double> tmp(n);
    auto put = [&](int f, const double *src, int stride, int comp) -> int {
        if(!c.s.f[f]) return SPHGPU_OK;
        if(!src) {
            hipLaunchKernelGGL(fill_double_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, c.stream, c.s.f[f], n, 0.0);
            return SPHGPU_OK;
        }
        for(int64_t i = 0; i < n; ++i) tmp[i] = src[i * stride + comp];
        SPHGPU_HIP(hipMemcpy(c.s.f[f], tmp.data(), n * sizeof(double), hipMemcpyHostToDevice));
        return SPHGPU_OK;
    };
    int rc = 0;
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    for(int k = 0; k < D && !rc; ++k) {
        rc = rc ? rc : put(F_X + k, h->pos, D, k);
        rc = rc ? rc : put(F_V + k, h->vel, D, k);
        rc = rc ? rc : put(F_VP + k, h->vel_p, D, k);
        rc = rc ? rc : put(F_A + k, h->acc, D, k);
        rc = rc ? rc : put(F_GD + k, h->grad_density, D, k);
        rc = rc ? rc : put(F_GP + k, h->grad_pressure, D, k);
        for(int l = 0; l < D && !rc; ++l) rc = put(F_GV + 3 * k + l, h->grad_velocity, D * D, k * D + l);
    }
    const double *sc[] = {h->mass, h->dens, h->pres, h->ene, h->ene_p, h->dene, h->sml, h->sound, h->balsara, h->alpha, h->gradh, h->phi};
    for(int f = F_MASS; f <= F_PHI && !rc; ++f) rc = put(f, sc[f - F_MASS], 1, 0);
    if(!rc && c.s.f[F_HINV]) rc = put(F_HINV, nullptr, 1, 0);
    if(rc) return rc;
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, c.stream, c.s.id, n);
    if(h->neighbor) {
        SPHGPU_HIP(hipMemcpy(c.s.nb, h->neighbor, n * sizeof(int), hipMemcpyHostToDevice));
    } else {
        hipLaunchKernelGGL(fill_int_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, c.stream, c.s.nb, n, 0);
    }
    // list slots beyond n stay empty
    hipLaunchKernelGGL(fill_int_kernel, dim3(grid_for(c.cap_n)), dim3(BLOCK), 0, c.stream, c.list_cnt, c.cap_n, 0);
    hipLaunchKernelGGL(fill_int_kernel, dim3(grid_for(c.cap_n)), dim3(BLOCK), 0, c.stream, c.list_cnt2, c.cap_n, 0);
    SPHGPU_HIP(hipGetLastError());
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    c.uploaded = true;
    c.first_pre = true;
    c.tree_ready = false;
    c.have_gather = c.have_force = false;
    return SPHGPU_OK;
}

int sphgpu_download(void *ptr, sphgpu_particles *h)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_DATA();
    if(!h || h->n != c.n) return set_error(SPHGPU_E_INVALID, "particle count mismatch");
    const int64_t n = c.n;
    const int D = c.D;
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    std::vector<int> id(n);
    SPHGPU_HIP(hipMemcpy(id.data(), c.s.id, n * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<double> tmp(n);
    auto get = [&](int f, double *dst, int stride, int comp) -> int {
        if(!dst) return SPHGPU_OK;
        if(!c.s.f[f]) {
            for(int64_t i = 0; i < n; ++i) dst[(int64_t)id[i] * stride + comp] = 0.0;
            return SPHGPU_OK;
        }
        SPHGPU_HIP(hipMemcpy(tmp.data(), c.s.f[f], n * sizeof(double), hipMemcpyDeviceToHost));
        for(int64_t i = 0; i < n; ++i) dst[(int64_t)id[i] * stride + comp] = tmp[i];
        return SPHGPU_OK;
    };
    int rc = 0;
    for(int k = 0; k < D && !rc; ++k) {
        rc = rc ? rc : get(F_X + k, h->pos, D, k);
        rc = rc ? rc : get(F_V + k, h->vel, D, k);
        rc = rc ? rc : get(F_VP + k, h->vel_p, D, k);
        rc = rc ? rc : get(F_A + k, h->acc, D, k);
        rc = rc ? rc : get(F_GD + k, h->grad_density, D, k);
        rc = rc ? rc : get(F_GP + k, h->grad_pressure, D, k);
        for(int l = 0; l < D && !rc; ++l) rc = get(F_GV + 3 * k + l, h->grad_velocity, D * D, k * D + l);
    }
    double *sc[] = {h->mass, h->dens, h->pres, h->ene, h->ene_p, h->dene, h->sml, h->sound, h->balsara, h->alpha, h->gradh, h->phi};
    for(int f = F_MASS; f <= F_PHI && !rc; ++f) rc = get(f, sc[f - F_MASS], 1, 0);
    if(rc) return rc;
    if(h->id)
        for(int64_t i = 0; i < n; ++i) h->id[id[i]] = id[i];
    if(h->neighbor) {
        std::vector<int> nb(n);
        SPHGPU_HIP(hipMemcpy(nb.data(), c.s.nb, n * sizeof(int), hipMemcpyDeviceToHost));
        for(int64_t i = 0; i < n; ++i) h->neighbor[id[i]] = nb[i];
    }
    return SPHGPU_OK;
}

int sphgpu_make_tree(void *ptr)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_DATA();
    int rc = stage_tree(c);
    return rc ? rc : check_flags(c, "make_tree");
}

int sphgpu_pre_interaction(void *ptr)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_TREE();
    int rc = stage_pre(c);
    if(rc) return rc;
    if((rc = check_flags(c, "pre_interaction"))) return rc;
    SPHGPU_HIP(hipMemcpyAsync(c.scal_host, c.scal, NSCALARS * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    c.h_per_v_sig = c.scal_host[S_HVS];
    return SPHGPU_OK;
}

int sphgpu_fluid_force(void *ptr)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_TREE();
    int rc = stage_force(c);
    return rc ? rc : check_flags(c, "fluid_force");
}

int sphgpu_gravity_force(void *ptr)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_TREE();
    int rc = gravity_compute(c);
    return rc ? rc : check_flags(c, "gravity_force");
}

int sphgpu_timestep(void *ptr, double *dt_out)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_DATA();
    int rc = step_timestep(c);
    if(rc) return rc;
    SPHGPU_HIP(hipMemcpyAsync(c.scal_host, c.scal, NSCALARS * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    c.dt = c.scal_host[S_DT];
    if(dt_out) *dt_out = c.dt;
    return SPHGPU_OK;
}

int sphgpu_predict(void *ptr)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_DATA();
    return step_predict(c);
}

int sphgpu_correct(void *ptr)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_DATA();
    return step_correct(c);
}

int sphgpu_initialize(void *ptr)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_DATA();
    int rc;
    if((rc = step_init_fields(c))) return rc;
    if((rc = sphgpu_make_tree(ptr))) return rc;
    if((rc = sphgpu_pre_interaction(ptr))) return rc;
    if((rc = sphgpu_fluid_force(ptr))) return rc;
    return sphgpu_gravity_force(ptr);
}

// Adds the spans of the pending timer sets (steps that have completed) to
// timer_ms, oldest first.
static void timers_flush(Ctx &c)
{
    for(; c.tset_pending > 0; --c.tset_pending) {
        const int t = (c.tset_cur - c.tset_pending + TIMER_SETS) % TIMER_SETS;
        hipEvent_t *ev = c.tev[t], *eg = c.tev[t] + 16;
        const bool grav = c.tset_grav[t];
        for(int k = 0; k < 7; ++k) {
            float ms = 0.0f;
            // with gravity on, its timer is the walk's and the pair pass's
            // spans on their stream, which overlap the other stages; the
            // wait for their end (ev[5] to ev[6]) is in no stage, and their
            // sums are added to acc in the corrector's launch
            if(k == 5 && grav) {
                float ms2 = 0.0f;
                (void)hipEventElapsedTime(&ms, eg[0], eg[1]);
                (void)hipEventElapsedTime(&ms2, eg[2], eg[3]);
                ms += ms2;
            } else {
                (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
            }
            c.timer_ms[k] += ms;
        }
        // kernel spans: search, pre-interaction pairs, force pairs
        for(int k = 0; k < 3; ++k) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[8 + 2 * k], ev[9 + 2 * k]);
            c.timer_ms[7 + k] += ms;
        }
        if(grav) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, eg[0], eg[1]);
            c.timer_ms[10] += ms;
            (void)hipEventElapsedTime(&ms, eg[2], eg[3]);
            c.timer_ms[11] += ms;
        }
    }
}

int sphgpu_integrate(void *ptr, double fixed_dt)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_TREE();
    int rc;
    if(c.timers_on) {
        // this step's timer set (the oldest pending ones are read first
        // when every set is in use)
        if(c.tset_pending == TIMER_SETS) timers_flush(c);
        c.ev = c.tev[c.tset_cur];
        c.ev_g = c.tev[c.tset_cur] + 16;
    }
    hipEvent_t *ev = c.ev;
    auto mark = [&](int k) {
        if(c.timers_on) (void)hipEventRecord(ev[k], c.stream);
    };
    c.kernel_marks = c.timers_on;
    mark(0);
    if(fixed_dt > 0.0) hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, c.stream, c.scal, (int)S_DT, fixed_dt);
    else if((rc = step_timestep(c))) return rc;
    mark(1);
    if((rc = step_predict(c))) return rc;
    mark(2);
    // the combined pre-interaction (not the first, initial-smoothing one)
    // rewrites h, 1/h, P, grad-h and the neighbour counts, and the force and
    // gravity passes acc, du/dt and phi, before any kernel reads them
    c.reorder_light = !c.first_pre;
    rc = stage_tree(c);
    c.reorder_light = false;
    if(rc) return rc;
    // Gravity runs on the second stream in two parts: the tree walk (vector-
    // ALU bound; needs the tree and positions only), started once the tree
    // is built, beside the neighbour search (which waits on dependent loads)
    // and the pre-interaction; then the recorded leaves' particle pairs
    // (need h), beside the force pass.  The gravity acceleration is added to
    // acc once both streams are done.
    const bool grav = c.p.gravity_is_valid;
    mark(3);
    c.speculate = true;
    if((rc = stage_pre(c, grav))) {
        c.speculate = false;
        c.flags_pending = false;
        // later work on the main stream must not overtake the walk
        if(grav) (void)hipStreamWaitEvent(c.stream, c.ev_grav, 0);
        return rc;
    }
    mark(4);
    // the leaf pairs start once the force pass has packed its records
    c.after_pack = grav ? launch_gravity_leaves : nullptr;
    rc = stage_force(c);
    c.after_pack = nullptr;
    c.speculate = false;
    c.flags_pending = false;
    if(rc) {
        if(grav) (void)hipStreamWaitEvent(c.stream, c.ev_grav, 0);
        return rc;
    }
    mark(5);
    if(grav) SPHGPU_HIP(hipStreamWaitEvent(c.stream, c.ev_grav, 0));
    mark(6);
    // the gravity sums are added to acc by the corrector's launch
    if((rc = step_correct(c, grav))) return rc;
    mark(7);
    c.kernel_marks = false;
    if((rc = check_flags(c, "integrate", true))) return rc;  // synchronises the stream
    c.dt = c.scal_host[S_DT];
    c.h_per_v_sig = c.scal_host[S_HVS];
    c.time += c.dt;
    ++c.steps;
    if(c.timers_on) {
        // read when the timers are (sphgpu_get_timers)
        c.tset_grav[c.tset_cur] = grav;
        c.tset_cur = (c.tset_cur + 1) % TIMER_SETS;
        ++c.tset_pending;
    }
    return SPHGPU_OK;
}

// One full step with a fixed dt, without timers (the rollout and its
// recomputation use exactly this sequence).
static int rollout_step(Ctx &c, double dt)
{
    int rc;
    hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, c.stream, c.scal, (int)S_DT, dt);
    if((rc = step_predict(c))) return rc;
    if((rc = stage_tree(c))) return rc;
    if((rc = stage_pre(c))) return rc;
    if((rc = stage_force(c))) return rc;
    if((rc = gravity_compute(c))) return rc;
    return step_correct(c);
}

// The force evaluation of step `slot - 1` (slot >= 1), or of
// Solver::initialize (slot 0), recomputed from its checkpoint.
static int rollout_recompute(Ctx &c, int slot)
{
    int rc;
    if((rc = adjoint_restore(c, slot))) return rc;
    if(slot == 0) {
        c.first_pre = true;
        if((rc = step_init_fields(c))) return rc;
    } else {
        c.first_pre = false;
        hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, c.stream, c.scal, (int)S_DT, c.ro.dt);
        if((rc = step_predict(c))) return rc;
    }
    if((rc = stage_tree(c))) return rc;
    if((rc = stage_pre(c))) return rc;
    return stage_force(c);
}

static int rollout_supported(const Ctx &c)
{
    const sphgpu_params &p = c.p;
    if(p.sph_type != SPHGPU_SSPH) return set_error(SPHGPU_E_INVALID, "rollout adjoint: SSPH only");
    if(!p.iterative_sml) return set_error(SPHGPU_E_INVALID, "rollout adjoint: needs iterativeSmoothingLength");
    if(p.av_use_time_dependent) return set_error(SPHGPU_E_INVALID, "rollout adjoint: time-dependent viscosity is not differentiated");
    if(p.ac_is_valid) return set_error(SPHGPU_E_INVALID, "rollout adjoint: artificial conductivity is not differentiated");
    if(p.gravity_is_valid) return set_error(SPHGPU_E_INVALID, "rollout adjoint: gravity is not differentiated");
    return SPHGPU_OK;
}

int sphgpu_rollout(void *ptr, int steps, double dt)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_DATA();
    int rc;
    if(steps < 1) return set_error(SPHGPU_E_INVALID, "rollout: steps must be >= 1");
    if(!(dt > 0.0)) return set_error(SPHGPU_E_INVALID, "rollout: dt must be > 0");
    if((rc = rollout_supported(c))) return rc;
    if((rc = adjoint_alloc(c, steps))) return rc;
    c.ro.dt = dt;
    c.ro.t0 = c.time;
    if((rc = adjoint_save(c, 0))) return rc;
    c.first_pre = true;
    if((rc = step_init_fields(c))) return rc;
    if((rc = stage_tree(c))) return rc;
    if((rc = stage_pre(c))) return rc;
    if((rc = stage_force(c))) return rc;
    if((rc = gravity_compute(c))) return rc;
    for(int k = 0; k < steps; ++k) {
        if((rc = adjoint_save(c, k + 1))) return rc;
        if((rc = rollout_step(c, dt))) return rc;
    }
    if((rc = adjoint_save(c, steps + 1))) return rc;
    SPHGPU_HIP(hipMemcpyAsync(c.scal_host, c.scal, NSCALARS * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if((rc = check_flags(c, "rollout"))) return rc;
    c.dt = dt;
    c.h_per_v_sig = c.scal_host[S_HVS];
    c.time = c.ro.t0 + steps * dt;
    c.steps += steps;
    c.ro.steps = steps;
    return SPHGPU_OK;
}

int sphgpu_rollout_adjoint(void *ptr, const double *dl_dx, const double *dl_dv, const double *dl_du, double *dx0,
                           double *dv0, double *du0)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_DATA();
    int rc;
    if(c.ro.steps < 1 || c.ro.n != c.n) return set_error(SPHGPU_E_STATE, "rollout adjoint: no stored rollout");
    const int K = c.ro.steps;
    if((rc = adjoint_seed(c, dl_dx, dl_dv, dl_du))) return rc;
    if((rc = ensure_lists(c))) return rc;
    // the recomputed steps expand their candidate tiles into lists, which
    // the adjoint passes read
    struct ListsOn {
        Ctx &c;
        explicit ListsOn(Ctx &c_) : c(c_) { c.need_lists = true; }
        ~ListsOn() { c.need_lists = false; }
    } lists_on(c);
    for(int k = K; k >= 0; --k) {
        if((rc = rollout_recompute(c, k))) return rc;
        if((rc = adjoint_step(c, k > 0))) return rc;
    }
    if((rc = check_flags(c, "rollout_adjoint"))) return rc;
    if((rc = adjoint_fetch(c, dx0, dv0, du0))) return rc;
    // leave the context at the final state of the rollout, with its tree
    if((rc = adjoint_restore(c, K + 1))) return rc;
    c.first_pre = false;
    if((rc = stage_tree(c))) return rc;
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    return SPHGPU_OK;
}

int sphgpu_get_time(void *ptr, double *time, double *dt, double *hvs)
{
    SPHGPU_CTX(ptr);
    if(time) *time = c.time;
    if(dt) *dt = c.dt;
    if(hvs) *hvs = c.h_per_v_sig;
    return SPHGPU_OK;
}

int sphgpu_set_time(void *ptr, double time, double dt)
{
    SPHGPU_CTX(ptr);
    c.time = time;
    c.dt = dt;
    hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, c.stream, c.scal, (int)S_DT, dt);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

// Resume from a state saved after a full step: the uploaded particles keep
// their h, density (the next guess), alpha, acceleration and du/dt; the
// one-off initial smoothing is skipped and the tree is built.
int sphgpu_restart(void *ptr, double time, double dt, double h_per_v_sig)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_DATA();
    if(!(h_per_v_sig > 0.0)) return set_error(SPHGPU_E_INVALID, "restart: h_per_v_sig must be > 0");
    c.time = time;
    c.dt = dt;
    c.h_per_v_sig = h_per_v_sig;
    hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, c.stream, c.scal, (int)S_DT, dt);
    hipLaunchKernelGGL(set_scalar_kernel, dim3(1), dim3(1), 0, c.stream, c.scal, (int)S_HVS, h_per_v_sig);
    SPHGPU_HIP(hipGetLastError());
    c.first_pre = false;
    int rc = stage_tree(c);
    return rc ? rc : check_flags(c, "restart");
}

int sphgpu_energy(void *ptr, double *out)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_DATA();
    if(!out) return set_error(SPHGPU_E_INVALID, "null output");
    return step_energy(c, out);
}

int sphgpu_get_stats(void *ptr, sphgpu_stats *st)
{
    SPHGPU_CTX(ptr);
    if(!st) return set_error(SPHGPU_E_INVALID, "null output");
    SPHGPU_HIP(hipMemcpyAsync(c.counters_host, c.counters, NCOUNTER_ENTRIES * sizeof(int64_t), hipMemcpyDeviceToHost,
                              c.stream));
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    // the rows of the statistics: sums, and the largest of the maxima
    int64_t k[STAT_ROW] = {};
    for(int r = 0; r < STAT_SLOTS; ++r) {
        const int64_t *row = c.counters_host + (size_t)r * STAT_ROW;
        for(int q = 0; q < NCOUNTERS; ++q) {
            if(q == C_MAXLIST || q == C_GRAV_MAXREC) k[q] = std::max(k[q], row[q]);
            else k[q] += row[q];
        }
    }
    st->steps = c.steps;
    st->gather_pairs = k[C_GATHER_PAIRS];
    st->force_pairs = k[C_FORCE_PAIRS];
    st->gravity_pp = k[C_GRAV_PP];
    st->gravity_pn = k[C_GRAV_PN];
    st->gravity_opened = k[C_GRAV_OPEN];
    st->search_nodes = k[C_SEARCH_NODES];
    st->search_candidates = k[C_SEARCH_CAND];
    st->gravity_wave_steps = k[C_GRAV_STEPS];
    st->gravity_rewalks = k[C_GRAV_REWALK];
    st->gravity_pp_slots = k[C_GRAV_SLOTS];
    st->gravity_max_records = k[C_GRAV_MAXREC];
    st->newton_failures = k[C_NEWTON_FAIL];
    st->newton_iterations = k[C_NEWTON_ITERS];
    st->search_truncations = k[C_TRUNC];
    st->list_overflows = k[C_OVERFLOW];
    st->max_list_length = k[C_MAXLIST];
    int nodes = 0;
    if(c.tree_ready) SPHGPU_HIP(hipMemcpy(&nodes, c.t.num_nodes, sizeof(int), hipMemcpyDeviceToHost));
    st->tree_nodes = nodes;
    st->tree_max_level = 0;
    if(c.tree_ready && nodes > 0) {
        std::vector<int4> rng(nodes);
        SPHGPU_HIP(hipMemcpy(rng.data(), c.t.rng, nodes * sizeof(int4), hipMemcpyDeviceToHost));
        int ml = 1;
        for(const int4 &r : rng) ml = std::max(ml, r.w & 255);
        st->tree_max_level = ml;
    }
    return SPHGPU_OK;
}

int sphgpu_get_timers(void *ptr, sphgpu_timers *t)
{
    SPHGPU_CTX(ptr);
    if(!t) return set_error(SPHGPU_E_INVALID, "null output");
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    if(c.gstream) SPHGPU_HIP(hipStreamSynchronize(c.gstream));
    timers_flush(c);
    t->timestep = c.timer_ms[0];
    t->predict = c.timer_ms[1];
    t->tree = c.timer_ms[2];
    t->pre_interaction = c.timer_ms[3];
    t->fluid_force = c.timer_ms[4];
    t->gravity = c.timer_ms[5];
    t->correct = c.timer_ms[6];
    t->search = c.timer_ms[7];
    t->pre_pairs = c.timer_ms[8];
    t->force_pairs = c.timer_ms[9];
    t->gravity_walk = c.timer_ms[10];
    t->gravity_leaf = c.timer_ms[11];
    return SPHGPU_OK;
}

int sphgpu_reset_stats(void *ptr)
{
    SPHGPU_CTX(ptr);
    c.trunc_seen = 0;
    hipLaunchKernelGGL(fill_i64_kernel, dim3(1), dim3(BLOCK), 0, c.stream, c.counters, NCOUNTER_ENTRIES);
    SPHGPU_HIP(hipGetLastError());
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    for(double &x : c.timer_ms) x = 0.0;
    c.tset_pending = 0;  // steps before the reset are not counted
    c.steps = 0;
    return SPHGPU_OK;
}

int sphgpu_enable_timers(void *ptr, int enable)
{
    SPHGPU_CTX(ptr);
    if(enable && !c.tev[0][0]) {
        for(int t = 0; t < TIMER_SETS; ++t)
            for(int k = 0; k < TIMER_EVENTS; ++k) SPHGPU_HIP(hipEventCreateWithFlags(&c.tev[t][k], EV_SCOPE));
        c.ev = c.tev[0];
        c.ev_g = c.tev[0] + 16;
    }
    c.timers_on = enable != 0;
    return SPHGPU_OK;
}

int sphgpu_enable_stats(void *ptr, int enable)
{
    SPHGPU_CTX(ptr);
    c.stats_on = enable != 0;
    return SPHGPU_OK;
}

int sphgpu_neighbor_sets(void *ptr, int kind, int32_t *counts, int32_t *lists, int64_t cap)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_DATA();
    if(!counts || !lists) return set_error(SPHGPU_E_INVALID, "null output");
    if((kind == 0 && !c.have_gather) || (kind != 0 && !c.have_force))
        return set_error(SPHGPU_E_STATE, "sets not recorded (enable sphgpu_set_record_sets before the stage)");
    const int64_t n = c.n;
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    std::vector<int> id(n), cnt(n);
    const size_t lsz = (size_t)c.cap_n * (size_t)c.cap;
    std::vector<int> lst(lsz);
    SPHGPU_HIP(hipMemcpy(id.data(), c.s.id, n * sizeof(int), hipMemcpyDeviceToHost));
    SPHGPU_HIP(hipMemcpy(cnt.data(), kind == 0 ? c.rec_gather_cnt : c.rec_force_cnt, n * sizeof(int), hipMemcpyDeviceToHost));
    SPHGPU_HIP(hipMemcpy(lst.data(), kind == 0 ? c.rec_gather : c.rec_force, lsz * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<std::vector<int>> sets(n);
    for(int64_t i = 0; i < n; ++i) {
        std::vector<int> &m = sets[id[i]];
        m.resize(cnt[i]);
        for(int k = 0; k < cnt[i]; ++k) m[k] = id[lst[list_at(i, k, c.cap)]];
        std::sort(m.begin(), m.end());
    }
    int64_t off = 0;
    for(int64_t i = 0; i < n; ++i) {
        counts[i] = (int32_t)sets[i].size();
        for(int j : sets[i]) {
            if(off >= cap) return set_error(SPHGPU_E_INVALID, "list buffer too small");
            lists[off++] = j;
        }
    }
    return SPHGPU_OK;
}

int sphgpu_tree_info(void *ptr, int64_t *num_nodes, int32_t *level, int32_t *count, double *edge, double *center, double *mass,
                     double *com, int64_t cap)
{
    SPHGPU_CTX(ptr);
    SPHGPU_NEED_TREE();
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    int nodes = 0;
    SPHGPU_HIP(hipMemcpy(&nodes, c.t.num_nodes, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int4> rng(nodes);
    std::vector<double4> geo(nodes), mom(nodes);
    SPHGPU_HIP(hipMemcpy(rng.data(), c.t.rng, nodes * sizeof(int4), hipMemcpyDeviceToHost));
    SPHGPU_HIP(hipMemcpy(geo.data(), c.t.geo, nodes * sizeof(double4), hipMemcpyDeviceToHost));
    SPHGPU_HIP(hipMemcpy(mom.data(), c.t.mom, nodes * sizeof(double4), hipMemcpyDeviceToHost));
    const int D = c.D;
    for(int64_t k = 0; k < nodes && k < cap; ++k) {
        if(level) level[k] = rng[k].w & 255;
        if(count) count[k] = rng[k].y - rng[k].x;
        if(edge) edge[k] = geo[k].w;
        const double cc[3] = {geo[k].x, geo[k].y, geo[k].z};
        const double cm[3] = {mom[k].x, mom[k].y, mom[k].z};
        for(int d = 0; d < D; ++d) {
            if(center) center[k * D + d] = cc[d];
            if(com) com[k * D + d] = cm[d];
        }
        if(mass) mass[k] = mom[k].w;
    }
    *num_nodes = nodes;
    return SPHGPU_OK;
}

int sphgpu_synchronize(void *ptr)
{
    SPHGPU_CTX(ptr);
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    return SPHGPU_OK;
}

}  // extern "C"
