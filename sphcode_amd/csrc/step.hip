// Streaming stages and global reductions: per-particle initialisation,
// TimeStep, predict/correct kick-drift and the energy sums.  Reductions are
// two-pass (block partials, then one block), in FP64, deterministic.
#include <mutex>

#include "ctx.hpp"

namespace sphgpu {

// Solver::initialize per-particle loop (src/solver.cpp:392-404).
template <int D> __global__ void __launch_bounds__(BLOCK) init_kernel(State s, int n, double alpha, double cs)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
    s.f[F_ALPHA][i] = alpha;
    s.f[F_BAL][i] = 1.0;
    s.f[F_SOUND][i] = sqrt(cs * s.f[F_ENE][i]);
}

// TimeStep::calculation (src/timestep.cpp:18-38), first pass: per-block
// min of c_force sqrt(h / |a|) over |a| > 0, grid-stride over the particles
// (at most 1024 blocks, one partial each).
template <int D>
__global__ void __launch_bounds__(BLOCK) dt_partial_kernel(const double *sml, const double *ax, const double *ay,
                                                           const double *az, int n, double cf, double *partial)
{
    double m = 1.7976931348623157e308;
    for(int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        double a2 = ax[i] * ax[i];
        if constexpr(D > 1) a2 += ay[i] * ay[i];
        if constexpr(D > 2) a2 += az[i] * az[i];
        const double a = sqrt(a2);
        if(a > 0.0) {
            const double d = cf * sqrt(sml[i] / a);
            m = d < m ? d : m;
        }
    }
    __shared__ double sh[BLOCK / WAVE];
    m = wave_minv(m);
    if(wave_lane() == 0) sh[threadIdx.x / WAVE] = m;
    __syncthreads();
    if(threadIdx.x == 0) {
        for(int w = 1; w < BLOCK / WAVE; ++w) m = sh[w] < m ? sh[w] : m;
        partial[blockIdx.x] = sh[0] < m ? sh[0] : m;
    }
}

// Second pass: dt = min(c_sound * min(h / v_sig), min dt_force).
__global__ void dt_final_kernel(const double *partial, int nb, double csound, const double *h_per_v_sig, double *dt)
{
    __shared__ double sh[BLOCK / WAVE];
    double m = 1.7976931348623157e308;
    for(int b = threadIdx.x; b < nb; b += blockDim.x) m = partial[b] < m ? partial[b] : m;
    m = wave_minv(m);
    if(wave_lane() == 0) sh[threadIdx.x / WAVE] = m;
    __syncthreads();
    if(threadIdx.x == 0) {
        for(int w = 1; w < (int)(blockDim.x / WAVE); ++w) m = sh[w] < m ? sh[w] : m;
        m = sh[0] < m ? sh[0] : m;
        const double ds = csound * *h_per_v_sig;
        *dt = ds < m ? ds : m;
    }
}

// The two passes over one state (grid of at most DT_MAX_BLOCKS blocks).
constexpr int DT_MAX_BLOCKS = 1024;
static void launch_dt(int D, int n, const double *sml, const double *const acc[3], const double *hvs, double cf,
                      double cs, double *partial, double *dt, hipStream_t st)
{
    const int nb = std::min(grid_for(n), DT_MAX_BLOCKS);
    switch(D) {
    case 1: hipLaunchKernelGGL(dt_partial_kernel<1>, dim3(nb), dim3(BLOCK), 0, st, sml, acc[0], acc[1], acc[2], n, cf, partial); break;
    case 2: hipLaunchKernelGGL(dt_partial_kernel<2>, dim3(nb), dim3(BLOCK), 0, st, sml, acc[0], acc[1], acc[2], n, cf, partial); break;
    default: hipLaunchKernelGGL(dt_partial_kernel<3>, dim3(nb), dim3(BLOCK), 0, st, sml, acc[0], acc[1], acc[2], n, cf, partial); break;
    }
    hipLaunchKernelGGL(dt_final_kernel, dim3(1), dim3(BLOCK), 0, st, partial, nb, cs, hvs, dt);
}

__global__ void min_final_kernel(const double *partial, int nb, double *scal, int slot)
{
    __shared__ double sh[BLOCK / WAVE];
    double m = 1.7976931348623157e308;
    for(int b = threadIdx.x; b < nb; b += blockDim.x) m = partial[b] < m ? partial[b] : m;
    m = wave_minv(m);
    if(wave_lane() == 0) sh[threadIdx.x / WAVE] = m;
    __syncthreads();
    if(threadIdx.x == 0) {
        for(int w = 1; w < (int)(blockDim.x / WAVE); ++w) m = sh[w] < m ? sh[w] : m;
        scal[slot] = sh[0] < m ? sh[0] : m;
    }
}

// Solver::predict (src/solver.cpp:431-456), operation order kept.
template <int D> __global__ void __launch_bounds__(BLOCK) predict_kernel(State s, int n, const double *scal, double cs, BoxDev box)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
    const double dt = scal[S_DT];
    const double hdt = 0.5 * dt;
    double x[D];
#pragma unroll
    for(int k = 0; k < D; ++k) {
        const double v = s.f[F_V + k][i], a = s.f[F_A + k][i];
        const double vp = v + a * hdt;
        s.f[F_VP + k][i] = vp;
        x[k] = s.f[F_X + k][i] + vp * dt;
        s.f[F_V + k][i] = v + a * dt;
    }
    const double u = s.f[F_ENE][i], du = s.f[F_DENE][i];
    s.f[F_ENEP][i] = u + du * hdt;
    const double un = u + du * dt;
    s.f[F_ENE][i] = un;
    s.f[F_SOUND][i] = sqrt(cs * un);
    wrap<D>(box, x);
#pragma unroll
    for(int k = 0; k < D; ++k) s.f[F_X + k][i] = x[k];
}

// Solver::correct (src/solver.cpp:458-474).
// With gacc, the gravity acceleration is first added to acc (as
// gravity_combine_kernel does), in the same launch.
template <int D> __global__ void __launch_bounds__(BLOCK) correct_kernel(State s, int n, const double *scal, double cs,
                                                                      const double *gacc, int64_t stride)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
    const double hdt = 0.5 * scal[S_DT];
#pragma unroll
    for(int k = 0; k < D; ++k) {
        double a = s.f[F_A + k][i];
        if(gacc) {
            a += gacc[k * stride + i];
            s.f[F_A + k][i] = a;
        }
        s.f[F_V + k][i] = s.f[F_VP + k][i] + a * hdt;
    }
    const double u = s.f[F_ENEP][i] + s.f[F_DENE][i] * hdt;
    s.f[F_ENE][i] = u;
    s.f[F_SOUND][i] = sqrt(cs * u);
}

// Energy sums of Output::output_energy (src/output.cpp:66-90) and momentum.
constexpr int NE = 7;  // kinetic, thermal, potential, momentum x3
template <int D> __global__ void __launch_bounds__(BLOCK) energy_partial_kernel(State s, int n, double *partial)
{
    double e[NE] = {0, 0, 0, 0, 0, 0, 0};
    for(int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        const double m = s.f[F_MASS][i];
        double v2 = 0.0;
#pragma unroll
        for(int k = 0; k < D; ++k) {
            const double v = s.f[F_V + k][i];
            v2 += v * v;
            e[3 + k] += m * v;
        }
        e[0] += 0.5 * m * v2;
        e[1] += m * s.f[F_ENE][i];
        e[2] += 0.5 * m * s.f[F_PHI][i];
    }
    __shared__ double sh[NE][BLOCK / WAVE];
#pragma unroll
    for(int k = 0; k < NE; ++k) {
        const double v = wave_sum(e[k]);
        if(wave_lane() == 0) sh[k][threadIdx.x / WAVE] = v;
    }
    __syncthreads();
    if(threadIdx.x < NE) {
        double t = 0.0;
        for(int w = 0; w < BLOCK / WAVE; ++w) t += sh[threadIdx.x][w];
        partial[blockIdx.x * NE + threadIdx.x] = t;
    }
}

__global__ void energy_final_kernel(const double *partial, int nb, double *out)
{
    if(threadIdx.x >= NE) return;
    double t = 0.0;
    for(int b = 0; b < nb; ++b) t += partial[b * NE + threadIdx.x];
    out[threadIdx.x] = t;
}

// FP32 mirrors of up to 16 FP64 fields.
struct MirrorArgs {
    const double *src[16];
    float *dst[16];
    int nf;
};

__global__ void __launch_bounds__(BLOCK) mirror_kernel(MirrorArgs m, int n)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
    for(int k = 0; k < m.nf; ++k) m.dst[k][i] = (float)m.src[k][i];
}

int mirror_fields(Ctx &c, const int *fields, int nfields)
{
    if(c.p.precision != SPHGPU_MIXED) return SPHGPU_OK;
    MirrorArgs m{};
    for(int k = 0; k < nfields; ++k) {
        const int f = fields[k];
        if(!c.s.f[f] || !c.s.g[f]) continue;
        m.src[m.nf] = c.s.f[f];
        m.dst[m.nf] = c.s.g[f];
        ++m.nf;
        if(m.nf == 16 || k == nfields - 1) {
            hipLaunchKernelGGL(mirror_kernel, dim3(grid_for(c.n)), dim3(BLOCK), 0, c.stream, m, (int)c.n);
            m.nf = 0;
        }
    }
    if(m.nf) hipLaunchKernelGGL(mirror_kernel, dim3(grid_for(c.n)), dim3(BLOCK), 0, c.stream, m, (int)c.n);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

// ------------------------------------------------------------------ host
BoxDev make_box(const Ctx &c)
{
    BoxDev box{};
    box.on = c.p.periodic_is_valid;
    for(int k = 0; k < 3; ++k) {
        box.lo[k] = c.p.range_min[k];
        box.hi[k] = c.p.range_max[k];
        box.len[k] = c.p.range_max[k] - c.p.range_min[k];
    }
    return box;
}

#define SPHGPU_DISPATCH_D(c, KERNEL, grid, ...)                                                     \
    switch((c).D) {                                                                                 \
    case 1: hipLaunchKernelGGL(KERNEL<1>, dim3(grid), dim3(BLOCK), 0, (c).stream, __VA_ARGS__); break; \
    case 2: hipLaunchKernelGGL(KERNEL<2>, dim3(grid), dim3(BLOCK), 0, (c).stream, __VA_ARGS__); break; \
    default: hipLaunchKernelGGL(KERNEL<3>, dim3(grid), dim3(BLOCK), 0, (c).stream, __VA_ARGS__); break; \
    }

int step_init_fields(Ctx &c)
{
    const double cs = c.p.gamma * (c.p.gamma - 1.0);
    SPHGPU_DISPATCH_D(c, init_kernel, grid_for(c.n), c.s, (int)c.n, c.p.av_alpha, cs);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

int reduce_min_partials(Ctx &c, int nblocks, int slot)
{
    hipLaunchKernelGGL(min_final_kernel, dim3(1), dim3(BLOCK), 0, c.stream, c.partial, nblocks, c.scal, slot);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

int step_timestep(Ctx &c)
{
    const double *acc[3] = {c.s.f[F_A], c.D > 1 ? c.s.f[F_A + 1] : nullptr, c.D > 2 ? c.s.f[F_A + 2] : nullptr};
    launch_dt(c.D, (int)c.n, c.s.f[F_SML], acc, c.scal + S_HVS, c.p.cfl_force, c.p.cfl_sound, c.partial,
              c.scal + S_DT, c.stream);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

int step_predict(Ctx &c)
{
    const double cs = c.p.gamma * (c.p.gamma - 1.0);
    const BoxDev box = make_box(c);
    SPHGPU_DISPATCH_D(c, predict_kernel, grid_for(c.n), c.s, (int)c.n, c.scal, cs, box);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

int step_correct(Ctx &c, bool add_gravity)
{
    const double cs = c.p.gamma * (c.p.gamma - 1.0);
    SPHGPU_DISPATCH_D(c, correct_kernel, grid_for(c.n), c.s, (int)c.n, c.scal, cs, add_gravity ? c.gacc : nullptr,
                      (int64_t)c.cap_n);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

int step_energy(Ctx &c, double *out)
{
    const int nb = std::min(grid_for(c.n), 1024);
    SPHGPU_DISPATCH_D(c, energy_partial_kernel, nb, c.s, (int)c.n, c.partial);
    double *dev_out = c.partial + 1024 * NE;
    hipLaunchKernelGGL(energy_final_kernel, dim3(1), dim3(WAVE), 0, c.stream, c.partial, nb, dev_out);
    SPHGPU_HIP(hipGetLastError());
    double host[NE];
    SPHGPU_HIP(hipMemcpyAsync(host, dev_out, sizeof(host), hipMemcpyDeviceToHost, c.stream));
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    out[0] = host[0];
    out[1] = host[1];
    out[2] = host[2];
    out[3] = host[0] + host[1] + host[2];
    for(int k = 0; k < c.D; ++k) out[4 + k] = host[3 + k];
    return SPHGPU_OK;
}

}  // namespace sphgpu

// ------------------------------------------------------------ no allocation
// sphgpu_timestep_noalloc keeps its block partials in this file-scope array
// (one copy per GPU, held until the process ends).  Host code never passes
// the array's name to a kernel: its device address is fetched once per
// device with hipGetSymbolAddress and kept in scratch_dev.
namespace sphgpu {

static __device__ __align__(256) double scratch[DT_MAX_BLOCKS];

constexpr int kMaxDevices = 64;
static double *scratch_dev[kMaxDevices] = {};
static std::mutex scratch_mu;

// Device address of this GPU's scratch array (the current device's), looked
// up on the first call on that device.
static int scratch_for_current_device(double **out)
{
    int dev = 0;
    SPHGPU_HIP(hipGetDevice(&dev));
    if(dev < 0 || dev >= kMaxDevices) return set_error(SPHGPU_E_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if(!scratch_dev[dev]) {
        double *p = nullptr;
        SPHGPU_HIP(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
        if(!p) return set_error(SPHGPU_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        scratch_dev[dev] = p;
    }
    *out = scratch_dev[dev];
    return SPHGPU_OK;
}

}  // namespace sphgpu

extern "C" int sphgpu_timestep_noalloc(int dim, int64_t n, const double *sml, const double *acc_x, const double *acc_y,
                                       const double *acc_z, const double *h_per_v_sig, double cfl_sound,
                                       double cfl_force, double *dt, void *stream)
{
    using namespace sphgpu;
    if(dim < 1 || dim > 3) return set_error(SPHGPU_E_INVALID, "sphgpu_timestep_noalloc: dim must be 1, 2 or 3");
    if(n < 1 || n > SPHGPU_NOALLOC_MAX_N)
        return set_error(SPHGPU_E_INVALID, "sphgpu_timestep_noalloc: n must be in [1, " +
                                               std::to_string((long long)SPHGPU_NOALLOC_MAX_N) + "], got " +
                                               std::to_string((long long)n));
    if(!sml || !acc_x || (dim > 1 && !acc_y) || (dim > 2 && !acc_z) || !h_per_v_sig || !dt)
        return set_error(SPHGPU_E_INVALID, "sphgpu_timestep_noalloc: null device pointer");
    double *partial = nullptr;
    int rc = scratch_for_current_device(&partial);
    if(rc) return rc;
    const double *acc[3] = {acc_x, dim > 1 ? acc_y : nullptr, dim > 2 ? acc_z : nullptr};
    launch_dt(dim, (int)n, sml, acc, h_per_v_sig, cfl_force, cfl_sound, partial, dt, (hipStream_t)stream);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}
