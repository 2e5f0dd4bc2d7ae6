// PreInteraction on the device, one thread per particle over its gather
// candidate list:
//   SSPH  src/pre_interaction.cpp:39-169 (Newton :227-283)
//   DISPH src/disph/d_pre_interaction.cpp:21-162 (Newton :174-230)
//   GSPH  src/gsph/g_pre_interaction.cpp:27-144
//
// Newton-Raphson for h solves rho(h) h^D = m N_ngb / A_D (SSPH, GSPH) or
// n(h) h^D = N_ngb / A_D (DISPH) over the candidates with r < h, starting
// from h_guess, and returns the iterate after the first step with
// |h - h_prev| < 1e-4 (h + h_prev); after ten steps it falls back to
// h_guess and counts a failure (the reference logs it).  Candidates beyond
// the search radius are never seen, as in the reference (a Newton h above
// the radius is counted as a truncation).
//
// One sweep over the final set r < h then accumulates the density sums,
// the signal velocity, and the velocity derivatives that the reference
// gathers in a second loop over the same set (Balsara switch and
// time-dependent viscosity, or the GSPH MUSCL gradients).
#include "ctx.hpp"

namespace sphgpu {

struct PreArgs {
    const int *list;
    const double *list_r;
    const int *cnt;
    int cap;
    double gamma;
    double nngb;
    double ratio;
    int iterative;
    int balsara;
    int tdav;
    double amax, amin, eps_av;
    const double *scal;   // S_DT
    int second_order;
    int64_t *counters;
    double *partial;      // per-block min of h / v_sig
};

template <int D> struct CurlAcc {
    double c[3] = {0.0, 0.0, 0.0};
    __device__ __forceinline__ void add(const double a[D], const double b[D], double m)
    {
        if constexpr(D == 3) {
            c[0] += (a[1] * b[2] - a[2] * b[1]) * m;
            c[1] += (a[2] * b[0] - a[0] * b[2]) * m;
            c[2] += (a[0] * b[1] - a[1] * b[0]) * m;
        } else if constexpr(D == 2) {
            c[2] += (a[0] * b[1] - a[1] * b[0]) * m;
        }
    }
    __device__ __forceinline__ double magnitude() const
    {
        if constexpr(D == 2) return fabs(c[2]);
        return sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    }
};

// Initial smoothing (src/pre_interaction.cpp:171-215): density with
// h = h_guess over candidates with r < h.
template <int D, int K>
__global__ void __launch_bounds__(BLOCK) pre_initial_kernel(State s, int n, PreArgs a)
{
    using Kern = typename KernelOf<D, K, double>::type;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
    const size_t base = ((size_t)i / WAVE) * a.cap * WAVE + i % WAVE;
    const int cnt = a.cnt[i];
    const double h = s.f[F_SML][i];
    const Kern ker(h);
    double rho = 0.0;
    for(int k = 0; k < cnt; ++k) {
        const double r = a.list_r[base + (size_t)k * WAVE];
        if(r < h) rho += s.f[F_MASS][a.list[base + (size_t)k * WAVE]] * ker.w(ker.q_of(r));
    }
    s.f[F_DENS][i] = rho;
}

template <int D, int K, int TYPE>
__global__ void __launch_bounds__(BLOCK) pre_kernel(State s, BoxDev box, int n, PreArgs a)
{
    using Kern = typename KernelOf<D, K, double>::type;
    constexpr bool NUMBER = TYPE == SPHGPU_DISPH;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const bool valid = i < n;
    double hvs = 1.7976931348623157e308;
    int64_t iters = 0, fails = 0, trunc = 0, pairs = 0;
    if(valid) {
        const size_t base = ((size_t)i / WAVE) * a.cap * WAVE + i % WAVE;
        const int cnt = a.cnt[i];
        const double mi = s.f[F_MASS][i];
        const double Hs = s.f[F_SML][i];  // search radius h_guess * ratio
        double h = Hs;
        if(a.iterative) {
            // PreInteraction::newton_raphson
            const double b = NUMBER ? a.nngb / UnitVolume<D>::value : mi * a.nngb / UnitVolume<D>::value;
            double hn = Hs / a.ratio;
            bool done = false;
            for(int it = 0; it < 10 && !done; ++it) {
                ++iters;
                const double hb = hn;
                const Kern ker(hn);
                double rho = 0.0, drho = 0.0;
                for(int k = 0; k < cnt; ++k) {
                    const double r = a.list_r[base + (size_t)k * WAVE];
                    if(r < hn) {
                        const double q = ker.q_of(r);
                        if constexpr(NUMBER) {
                            rho += ker.w(q);
                            drho += ker.dhw(q);
                        } else {
                            const double mj = s.f[F_MASS][a.list[base + (size_t)k * WAVE]];
                            rho += mj * ker.w(q);
                            drho += mj * ker.dhw(q);
                        }
                    }
                }
                const double f = rho * pow_dim<D>(hn) - b;
                const double df = drho * pow_dim<D>(hn) + D * rho * pow_dim1<D>(hn);
                hn -= f / df;
                if(fabs(hn - hb) < (hn + hb) * 1e-4) done = true;
            }
            if(done) {
                h = hn;
            } else {
                h = Hs / a.ratio;
                ++fails;
            }
            if(h > Hs) ++trunc;
        }

        // final sums over r < h
        double xi[D], vi[D];
#pragma unroll
        for(int k = 0; k < D; ++k) { xi[k] = s.f[F_X + k][i]; vi[k] = s.f[F_V + k][i]; }
        const double ui = s.f[F_ENE][i];
        const double ci = s.f[F_SOUND][i];
        const Kern ker(h);
        double rho = 0.0, drho = 0.0, nd = 0.0, dnd = 0.0, pb = 0.0, dpb = 0.0;
        double vsig_max = ci * 2.0;
        double divv = 0.0;
        CurlAcc<D> rot;
        double gd[D], gu[D], gv[D][D];
#pragma unroll
        for(int k = 0; k < D; ++k) {
            gd[k] = 0.0;
            gu[k] = 0.0;
#pragma unroll
            for(int l = 0; l < D; ++l) gv[k][l] = 0.0;
        }
        const bool need_dv = (TYPE != SPHGPU_GSPH && (a.balsara || a.tdav)) || (TYPE == SPHGPU_GSPH && a.second_order);
        int count = 0;
        for(int k = 0; k < cnt; ++k) {
            const double r = a.list_r[base + (size_t)k * WAVE];
            if(!(r < h)) continue;
            const int j = a.list[base + (size_t)k * WAVE];
            ++count;
            const double mj = s.f[F_MASS][j];
            const double q = ker.q_of(r);
            const double w = ker.w(q);
            if constexpr(TYPE == SPHGPU_DISPH) {
                const double dh = ker.dhw(q);
                const double uj = s.f[F_ENE][j];
                rho += mj * w;
                nd += w;
                pb += mj * uj * w;
                dpb += mj * uj * dh;
                dnd += dh;
            } else if constexpr(TYPE == SPHGPU_SSPH) {
                rho += mj * w;
                drho += mj * ker.dhw(q);
            } else {
                rho += mj * w;
            }
            double rij[D], vij[D];
            if(j != i || need_dv) {
                double xj[D];
#pragma unroll
                for(int l = 0; l < D; ++l) xj[l] = s.f[F_X + l][j];
                sep2<D>(box, xi, xj, rij);
#pragma unroll
                for(int l = 0; l < D; ++l) vij[l] = vi[l] - s.f[F_V + l][j];
            }
            if(j != i) {
                double rv = rij[0] * vij[0];
#pragma unroll
                for(int l = 1; l < D; ++l) rv += rij[l] * vij[l];
                const double vs = ci + s.f[F_SOUND][j] - 3.0 * rv / r;
                vsig_max = vs > vsig_max ? vs : vsig_max;
            }
            if(need_dv) {
                const double cw = ker.dw_coef(q, r);
                double dw[D];
#pragma unroll
                for(int l = 0; l < D; ++l) dw[l] = rij[l] * cw;
                if constexpr(TYPE == SPHGPU_GSPH) {
                    const double du = mj * (s.f[F_ENE][j] - ui);
#pragma unroll
                    for(int l = 0; l < D; ++l) {
                        gd[l] += dw[l] * mj;
                        gu[l] += dw[l] * du;
                    }
#pragma unroll
                    for(int c = 0; c < D; ++c) {
                        const double dvc = mj * (-vij[c]);
#pragma unroll
                        for(int l = 0; l < D; ++l) gv[c][l] += dw[l] * dvc;
                    }
                } else {
                    const double wgt = TYPE == SPHGPU_DISPH ? mj * s.f[F_ENE][j] : mj;
                    double vd = vij[0] * dw[0];
#pragma unroll
                    for(int l = 1; l < D; ++l) vd += vij[l] * dw[l];
                    divv -= wgt * vd;
                    rot.add(vij, dw, wgt);
                }
            }
        }
        pairs = count;
        const double gm1 = a.gamma - 1.0;
        s.f[F_SML][i] = h;
        s.f[F_HINV][i] = 1.0 / h;
        s.f[F_DENS][i] = rho;
        double pres;
        if constexpr(TYPE == SPHGPU_DISPH) {
            pres = gm1 * pb;
            s.f[F_GRADH][i] = h / (D * nd) * dpb / (1.0 + h / (D * nd) * dnd);
        } else {
            pres = gm1 * rho * ui;
            if constexpr(TYPE == SPHGPU_SSPH) s.f[F_GRADH][i] = 1.0 / (1.0 + h / (D * rho) * drho);
        }
        s.f[F_PRES][i] = pres;
        s.nb[i] = count;
        hvs = h / vsig_max;

        if constexpr(TYPE == SPHGPU_GSPH) {
            if(a.second_order) {
                const double rinv = 1.0 / rho;
#pragma unroll
                for(int l = 0; l < D; ++l) {
                    s.f[F_GD + l][i] = gd[l];
                    s.f[F_GP + l][i] = (gd[l] * ui + gu[l]) * gm1;
                }
#pragma unroll
                for(int c = 0; c < D; ++c)
#pragma unroll
                    for(int l = 0; l < D; ++l) s.f[F_GV + 3 * c + l][i] = gv[c][l] * rinv;
            }
        } else {
            const double dt = a.scal[S_DT];
            double alpha = s.f[F_ALPHA][i];
            if(a.balsara && D != 1) {
                double rm;
                if constexpr(TYPE == SPHGPU_DISPH) {
                    const double p_inv = gm1 / pres;
                    divv *= p_inv;
                    rm = rot.magnitude() * fabs(p_inv);
                } else {
                    divv /= rho;
                    rm = rot.magnitude() / rho;
                }
                s.f[F_BAL][i] = fabs(divv) / (fabs(divv) + rm + 1e-4 * ci / h);
                if(a.tdav) {
                    const double tau_inv = a.eps_av * ci / h;
                    const double da = (-(alpha - a.amin) * tau_inv + fmax(-divv, 0.0) * (a.amax - alpha)) * dt;
                    s.f[F_ALPHA][i] = alpha + da;
                }
            } else if(a.tdav) {
                if constexpr(TYPE == SPHGPU_DISPH) divv *= gm1 / pres;
                else divv /= rho;
                const double tau_inv = a.eps_av * ci / h;
                const double si = fmax(-divv, 0.0);
                s.f[F_ALPHA][i] = (alpha + dt * tau_inv * a.amin + si * dt * a.amax) / (1.0 + dt * tau_inv + si * dt);
            }
        }
    }
    // reductions: min h/v_sig per block, counters per wave
    __shared__ double sh[BLOCK / WAVE];
    const double wmin = wave_minv(hvs);
    iters = wave_sum(iters);
    fails = wave_sum(fails);
    trunc = wave_sum(trunc);
    pairs = wave_sum(pairs);
    if(wave_lane() == 0) {
        sh[threadIdx.x / WAVE] = wmin;
        if(iters) atomicAdd((unsigned long long *)&a.counters[C_NEWTON_ITERS], (unsigned long long)iters);
        if(fails) atomicAdd((unsigned long long *)&a.counters[C_NEWTON_FAIL], (unsigned long long)fails);
        if(trunc) atomicAdd((unsigned long long *)&a.counters[C_TRUNC], (unsigned long long)trunc);
        if(pairs) atomicAdd((unsigned long long *)&a.counters[C_GATHER_PAIRS], (unsigned long long)pairs);
    }
    __syncthreads();
    if(threadIdx.x == 0) {
        double m = sh[0];
        for(int w = 1; w < BLOCK / WAVE; ++w) m = sh[w] < m ? sh[w] : m;
        a.partial[blockIdx.x] = m;
    }
}

static PreArgs make_args(Ctx &c)
{
    PreArgs a{};
    a.list = c.list;
    a.list_r = c.list_r;
    a.cnt = c.list_cnt;
    a.cap = c.cap;
    a.gamma = c.p.gamma;
    a.nngb = (double)c.p.neighbor_number;
    a.ratio = c.p.iterative_sml ? 1.2 : 1.0;
    a.iterative = c.p.iterative_sml;
    a.balsara = c.p.av_use_balsara_switch;
    a.tdav = c.p.av_use_time_dependent;
    a.amax = c.p.av_alpha_max;
    a.amin = c.p.av_alpha_min;
    a.eps_av = c.p.av_epsilon;
    a.scal = c.scal;
    a.second_order = c.p.gsph_is_2nd_order;
    a.counters = c.counters;
    a.partial = c.partial;
    return a;
}

template <int D, int K> static void launch_initial(Ctx &c, const PreArgs &a)
{
    hipLaunchKernelGGL((pre_initial_kernel<D, K>), dim3(grid_for(c.n)), dim3(BLOCK), 0, c.stream, c.s, (int)c.n, a);
}

int pre_initial(Ctx &c)
{
    const PreArgs a = make_args(c);
    const int K = c.p.kernel;
    switch(c.D) {
    case 1: K ? launch_initial<1, 1>(c, a) : launch_initial<1, 0>(c, a); break;
    case 2: K ? launch_initial<2, 1>(c, a) : launch_initial<2, 0>(c, a); break;
    default: K ? launch_initial<3, 1>(c, a) : launch_initial<3, 0>(c, a); break;
    }
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

template <int D, int K, int TYPE> static void launch_pre(Ctx &c, const PreArgs &a, const BoxDev &box)
{
    hipLaunchKernelGGL((pre_kernel<D, K, TYPE>), dim3(grid_for(c.n)), dim3(BLOCK), 0, c.stream, c.s, box, (int)c.n, a);
}

template <int D, int K> static void launch_pre_t(Ctx &c, const PreArgs &a, const BoxDev &box)
{
    switch(c.p.sph_type) {
    case SPHGPU_SSPH: launch_pre<D, K, SPHGPU_SSPH>(c, a, box); break;
    case SPHGPU_DISPH: launch_pre<D, K, SPHGPU_DISPH>(c, a, box); break;
    default: launch_pre<D, K, SPHGPU_GSPH>(c, a, box); break;
    }
}

BoxDev make_box(const Ctx &c);

int pre_compute(Ctx &c)
{
    const PreArgs a = make_args(c);
    const BoxDev box = make_box(c);
    const int K = c.p.kernel;
    switch(c.D) {
    case 1: K ? launch_pre_t<1, 1>(c, a, box) : launch_pre_t<1, 0>(c, a, box); break;
    case 2: K ? launch_pre_t<2, 1>(c, a, box) : launch_pre_t<2, 0>(c, a, box); break;
    default: K ? launch_pre_t<3, 1>(c, a, box) : launch_pre_t<3, 0>(c, a, box); break;
    }
    SPHGPU_HIP(hipGetLastError());
    return reduce_min_partials(c, grid_for(c.n), S_HVS);
}

}  // namespace sphgpu
