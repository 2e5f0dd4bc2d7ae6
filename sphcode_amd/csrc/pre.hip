// PreInteraction on the device, one wave per search group of at most 64
// particles, one lane per particle, over the group's candidate tiles: each
// tile's 64 neighbour records are staged once in LDS and every lane takes
// the members of its gather candidate set (front hit mask) among them:
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
//
// T is the pair arithmetic: double, or float for the mixed mode, where
// pair terms are evaluated in FP32 relative to particle i (separations
// taken in FP64, then rounded) and accumulated in FP64 (Acc in ctx.hpp),
// and the Newton update, its convergence test and every stored result
// stay FP64.
#include "tiles.hpp"

namespace sphgpu {

struct PreArgs {
    const int2 *groups;
    int ngroups;
    const int *cand;      // candidate tiles (Ctx::tile_cand)
    const uint64_t *mf;   // front (gather candidate) masks
    const int *gnc;
    int wcap;
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
    int stats;            // add to the statistics counters (errors are always counted)
    double *partial;      // per-block min of h / v_sig
    const void *prec;     // PreRec<T> per particle
    const double4 *pm;    // FP64 position and mass per particle (Newton sweeps)
    void *gpart;          // GravPart<T> per particle, written with h (gravity on)
};

constexpr int PSW = BLOCK / WAVE;  // waves (groups) per block

// |r_ij| of a gather candidate in the pair type from the FP64 r^2: the
// correctly rounded FP64 square root, or its FP32 form (gather_r_f32)
template <class T> __device__ __forceinline__ T gather_r(double r2)
{
    if constexpr(sizeof(T) == 8) return sqrt(r2);
    else return gather_r_f32(r2);
}

// Records read for j by the final sweep (FP32 mirrors in mixed mode), and
// the 32-byte position and mass records the Newton sweeps read.
template <int D, class T> __global__ void __launch_bounds__(BLOCK) pack_pre_kernel(State s, int n, PreRec<T> *out, double4 *pm)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
    PreRec<T> r{};
#pragma unroll
    for(int k = 0; k < 3; ++k) r.x[k] = k < D ? s.f[F_X + k][i] : 0.0;
    pm[i] = make_double4(r.x[0], r.x[1], r.x[2], s.f[F_MASS][i]);
    r.m = fld<T>(s, F_MASS, i);
    r.u = fld<T>(s, F_ENE, i);
    r.c = fld<T>(s, F_SOUND, i);
#pragma unroll
    for(int k = 0; k < D; ++k) r.v[k] = fld<T>(s, F_V + k, i);
    out[i] = r;
}

// Writes an FP64 result and, in mixed mode, its FP32 mirror.
template <class T> __device__ __forceinline__ void put(const State &s, int f, int i, double v)
{
    s.f[f][i] = v;
    if constexpr(sizeof(T) == 4) s.g[f][i] = (float)v;
}

template <int D, class T> struct CurlAcc {
    Acc<T> c[3];
    __device__ __forceinline__ void add(const T a[D], const T b[D], T m)
    {
        if constexpr(D == 3) {
            c[0].add((a[1] * b[2] - a[2] * b[1]) * m);
            c[1].add((a[2] * b[0] - a[0] * b[2]) * m);
            c[2].add((a[0] * b[1] - a[1] * b[0]) * m);
        } else if constexpr(D == 2) {
            c[2].add((a[0] * b[1] - a[1] * b[0]) * m);
        }
    }
    __device__ __forceinline__ double magnitude() const
    {
        if constexpr(D == 2) return fabs(c[2].val());
        const double x = c[0].val(), y = c[1].val(), z = c[2].val();
        return sqrt(x * x + y * y + z * z);
    }
};

// Initial smoothing (src/pre_interaction.cpp:171-215): density with
// h = h_guess over candidates with r < h.
template <class T> struct PosMass {
    double x[3];
    T m;
};

template <int D, int K, class T>
__global__ void __launch_bounds__(BLOCK) pre_initial_kernel(State s, BoxDev box, PreArgs a)
{
    using Kern = typename KernelOf<D, K, T>::type;
    __shared__ PosMass<T> l_r[PSW][WAVE];
    const int w = threadIdx.x / WAVE;
    const int lane = wave_lane();
    const int grp = xcd_block() * PSW + w;
    if(grp >= a.ngroups) return;
    const int2 gr = a.groups[grp];
    const int i = gr.x + lane;
    const bool valid = i < gr.y;
    const int ii = valid ? i : gr.x;
    double xi[D];
#pragma unroll
    for(int k = 0; k < D; ++k) xi[k] = s.f[F_X + k][ii];
    const T h = (T)s.f[F_SML][ii];
    const Kern ker(h);
    Acc<T> rho;
    const size_t gbase = (size_t)grp * a.wcap;
    const int ntile = (a.gnc[grp] + WAVE - 1) / WAVE;
    auto stage = [&](int tt) -> uint64_t {
        const size_t o = gbase + (size_t)tt * WAVE;
        const int jl = stream_ld(a.cand + o + lane);
        if(jl >= 0) {
#pragma unroll
            for(int k = 0; k < D; ++k) l_r[w][lane].x[k] = s.f[F_X + k][jl];
            l_r[w][lane].m = fld<T>(s, F_MASS, jl);
        }
        return valid ? stream_ld(a.mf + o + lane) : 0ull;
    };
    auto pair = [&](int kc) {
        double xj[D], d[D];
#pragma unroll
        for(int k = 0; k < D; ++k) xj[k] = l_r[w][kc].x[k];
        const T r = gather_r<T>(sep2<D>(box, xi, xj, d));
        if(r < h) rho.add(l_r[w][kc].m * ker.w(ker.q_of(r)));
    };
    tile_sweep(ntile, stage, pair);
    if(valid) s.f[F_DENS][i] = rho.val();
}

template <int D, int K, int TYPE, class T>
__global__ void __launch_bounds__(BLOCK) pre_kernel(State s, BoxDev box, int n, PreArgs a)
{
    using Kern = typename KernelOf<D, K, T>::type;
    constexpr bool NUMBER = TYPE == SPHGPU_DISPH;
    __shared__ PreRec<T> l_r[PSW][WAVE];
    __shared__ int l_j[PSW][WAVE];
    const int w = threadIdx.x / WAVE;
    // the Newton sweeps stage 32-byte position and mass records in the same
    // LDS as the final sweep's records (64 of them per wave)
    double4 *const l_pm = reinterpret_cast<double4 *>(&l_r[w][0]);
    static_assert(sizeof(PreRec<T>) >= sizeof(double4) && (WAVE * sizeof(PreRec<T>)) % alignof(double4) == 0,
                  "position and mass records fit the final sweep's slot");
    const int lane = wave_lane();
    // waves past the last group take part in the block's reduction only
    const int grp = xcd_block() * PSW + w;
    const bool have = grp < a.ngroups;
    const int2 gr = have ? a.groups[grp] : make_int2(0, 0);
    const int i = gr.x + lane;
    const bool valid = have && i < gr.y;
    const int ii = valid ? i : gr.x;
    const size_t gbase = (size_t)(have ? grp : 0) * a.wcap;
    const int ntile = have ? (a.gnc[grp] + WAVE - 1) / WAVE : 0;
    const PreRec<T> *prec = (const PreRec<T> *)a.prec;
    double hvs = 1.7976931348623157e308;
    int64_t iters = 0, fails = 0, trunc = 0, pairs = 0;
    if(have) {
        double xi[D];
#pragma unroll
        for(int k = 0; k < D; ++k) xi[k] = s.f[F_X + k][ii];
        const double mi = s.f[F_MASS][ii];
        const double Hs = s.f[F_SML][ii];  // search radius h_guess * ratio
        double h = Hs;
        if(a.iterative) {
            // PreInteraction::newton_raphson, every lane on its own iterate;
            // the wave sweeps the tiles once per iteration for the lanes
            // still iterating
            const double b = NUMBER ? a.nngb / UnitVolume<D>::value : mi * a.nngb / UnitVolume<D>::value;
            double hn = Hs / a.ratio;
            bool done = !valid;
            int it = 0;
            for(; it < 10; ++it) {
                if(ballot(!done) == 0) break;
                const double hb = hn;
                const T ht = (T)hn;
                const Kern ker(ht);
                Acc<T> rho, drho;
                auto nstage = [&](int tt) -> uint64_t {
                    const size_t o = gbase + (size_t)tt * WAVE;
                    const int jl = stream_ld(a.cand + o + lane);
                    if(jl >= 0) l_pm[lane] = a.pm[jl];
                    return done ? 0ull : stream_ld(a.mf + o + lane);
                };
                auto npair = [&](int kc) {
                    const double4 pj = l_pm[kc];
                    double xj[D], d[D];
                    xj[0] = pj.x;
                    if constexpr(D > 1) xj[1] = pj.y;
                    if constexpr(D > 2) xj[2] = pj.z;
                    const T r = gather_r<T>(sep2<D>(box, xi, xj, d));
                    if(r < ht) {
                        const T q = ker.q_of(r);
                        if constexpr(NUMBER) {
                            rho.add(ker.w(q));
                            drho.add(ker.dhw(q));
                        } else {
                            const T mj = (T)pj.w;
                            rho.add(mj * ker.w(q));
                            drho.add(mj * ker.dhw(q));
                        }
                    }
                };
                tile_sweep(ntile, nstage, npair);
                if(!done) {
                    ++iters;
                    const double f = rho.val() * pow_dim<D>(hn) - b;
                    const double df = drho.val() * pow_dim<D>(hn) + D * rho.val() * pow_dim1<D>(hn);
                    hn -= f / df;
                    if(fabs(hn - hb) < (hn + hb) * 1e-4) done = true;
                }
            }
            if(valid) {
                if(done) {
                    h = hn;
                } else {
                    h = Hs / a.ratio;
                    ++fails;
                }
                if(h > Hs) ++trunc;
            }
        }

        // final sums over r < h
        T vi[D];
#pragma unroll
        for(int k = 0; k < D; ++k) vi[k] = (T)s.f[F_V + k][ii];
        const double ui = s.f[F_ENE][ii];
        const double ci = s.f[F_SOUND][ii];
        const T ht = (T)h;
        const Kern ker(ht);
        Acc<T> rho, drho, nd, dnd, pb, dpb, divv;
        T vsig_max = (T)(ci * 2.0);
        CurlAcc<D, T> rot;
        Acc<T> gd[D], gu[D], gv[D][D];
        const bool need_dv = (TYPE != SPHGPU_GSPH && (a.balsara || a.tdav)) || (TYPE == SPHGPU_GSPH && a.second_order);
        const T ci_t = (T)ci, ui_t = (T)ui;
        int count = 0;
        auto fstage = [&](int tt) -> uint64_t {
            const size_t o = gbase + (size_t)tt * WAVE;
            const int jl = stream_ld(a.cand + o + lane);
            if(jl >= 0) l_r[w][lane] = prec[jl];
            l_j[w][lane] = jl;
            return valid ? stream_ld(a.mf + o + lane) : 0ull;
        };
        auto fpair = [&](int kc) {
            const PreRec<T> pj = l_r[w][kc];
            double xj[D], d[D];
#pragma unroll
            for(int l = 0; l < D; ++l) xj[l] = pj.x[l];
            const T r = gather_r<T>(sep2<D>(box, xi, xj, d));
            if(!(r < ht)) return;
            const int j = l_j[w][kc];
            ++count;
            const T mj = pj.m;
            const T q = ker.q_of(r);
            const T wk = ker.w(q);
            if constexpr(TYPE == SPHGPU_DISPH) {
                const T dh = ker.dhw(q);
                const T uj = pj.u;
                rho.add(mj * wk);
                nd.add(wk);
                pb.add(mj * uj * wk);
                dpb.add(mj * uj * dh);
                dnd.add(dh);
            } else if constexpr(TYPE == SPHGPU_SSPH) {
                rho.add(mj * wk);
                drho.add(mj * ker.dhw(q));
            } else {
                rho.add(mj * wk);
            }
            T rij[D], vij[D];
#pragma unroll
            for(int l = 0; l < D; ++l) {
                rij[l] = (T)d[l];
                vij[l] = vi[l] - pj.v[l];
            }
            if(j != i) {
                T rv = rij[0] * vij[0];
#pragma unroll
                for(int l = 1; l < D; ++l) rv += rij[l] * vij[l];
                const T vs = ci_t + pj.c - T(3) * rv / r;
                vsig_max = vs > vsig_max ? vs : vsig_max;
            }
            if(need_dv) {
                const T cw = ker.dw_coef(q, r);
                T dw[D];
#pragma unroll
                for(int l = 0; l < D; ++l) dw[l] = rij[l] * cw;
                if constexpr(TYPE == SPHGPU_GSPH) {
                    const T du = mj * (pj.u - ui_t);
#pragma unroll
                    for(int l = 0; l < D; ++l) {
                        gd[l].add(dw[l] * mj);
                        gu[l].add(dw[l] * du);
                    }
#pragma unroll
                    for(int c = 0; c < D; ++c) {
                        const T dvc = mj * (-vij[c]);
#pragma unroll
                        for(int l = 0; l < D; ++l) gv[c][l].add(dw[l] * dvc);
                    }
                } else {
                    const T wgt = TYPE == SPHGPU_DISPH ? mj * pj.u : mj;
                    T vd = vij[0] * dw[0];
#pragma unroll
                    for(int l = 1; l < D; ++l) vd += vij[l] * dw[l];
                    divv.add(-(wgt * vd));
                    rot.add(vij, dw, wgt);
                }
            }
        };
        // the cut-off r < h of the final sums, tested ahead
        auto fkeep = [&](int kc) -> bool {
            double xj[D], d[D];
#pragma unroll
            for(int l = 0; l < D; ++l) xj[l] = l_r[w][kc].x[l];
            return gather_r<T>(sep2<D>(box, xi, xj, d)) < ht;
        };
        // refined in 3D except for GSPH (measured slower on the 3D Sedov
        // case), see tiles.hpp
        if constexpr(D == 3 && TYPE != SPHGPU_GSPH) {
            auto frefine = [&](uint64_t m) { return refine_bits(m, fkeep); };
            tile_sweep(ntile, fstage, frefine, fpair);
        } else {
            tile_sweep(ntile, fstage, fpair);
        }
        if(valid) {
            pairs = count;
            const double gm1 = a.gamma - 1.0;
            const double rho_d = rho.val();
            put<T>(s, F_SML, i, h);
            put<T>(s, F_HINV, i, 1.0 / h);
            if(a.gpart) {
                // the gravity leaf pass's record of i, as pack_parts_kernel writes it
                GravPart<T> gp{};
#pragma unroll
                for(int k = 0; k < 3; ++k) gp.x[k] = k < D ? xi[k] : 0.0;
                gp.m = (T)mi;
                gp.einv = T(2) * (T)(1.0 / h);  // 2 x the FP32 mirror of 1/h in mixed mode
                ((GravPart<T> *)a.gpart)[i] = gp;
            }
            put<T>(s, F_DENS, i, rho_d);
            double pres;
            if constexpr(TYPE == SPHGPU_DISPH) {
                pres = gm1 * pb.val();
                const double nv = nd.val();
                put<T>(s, F_GRADH, i, h / (D * nv) * dpb.val() / (1.0 + h / (D * nv) * dnd.val()));
            } else {
                pres = gm1 * rho_d * ui;
                if constexpr(TYPE == SPHGPU_SSPH) put<T>(s, F_GRADH, i, 1.0 / (1.0 + h / (D * rho_d) * drho.val()));
            }
            put<T>(s, F_PRES, i, pres);
            s.nb[i] = count;
            hvs = h / (double)vsig_max;

            if constexpr(TYPE == SPHGPU_GSPH) {
                if(a.second_order) {
                    const double rinv = 1.0 / rho_d;
#pragma unroll
                    for(int l = 0; l < D; ++l) {
                        put<T>(s, F_GD + l, i, gd[l].val());
                        put<T>(s, F_GP + l, i, (gd[l].val() * ui + gu[l].val()) * gm1);
                    }
#pragma unroll
                    for(int c = 0; c < D; ++c)
#pragma unroll
                        for(int l = 0; l < D; ++l) put<T>(s, F_GV + 3 * c + l, i, gv[c][l].val() * rinv);
                }
            } else {
                const double dt = a.scal[S_DT];
                const double alpha = s.f[F_ALPHA][i];
                double dv = divv.val();
                if(a.balsara && D != 1) {
                    double rm;
                    if constexpr(TYPE == SPHGPU_DISPH) {
                        const double p_inv = gm1 / pres;
                        dv *= p_inv;
                        rm = rot.magnitude() * fabs(p_inv);
                    } else {
                        dv /= rho_d;
                        rm = rot.magnitude() / rho_d;
                    }
                    put<T>(s, F_BAL, i, fabs(dv) / (fabs(dv) + rm + 1e-4 * ci / h));
                    if(a.tdav) {
                        const double tau_inv = a.eps_av * ci / h;
                        const double da = (-(alpha - a.amin) * tau_inv + fmax(-dv, 0.0) * (a.amax - alpha)) * dt;
                        put<T>(s, F_ALPHA, i, alpha + da);
                    }
                } else if(a.tdav) {
                    if constexpr(TYPE == SPHGPU_DISPH) dv *= gm1 / pres;
                    else dv /= rho_d;
                    const double tau_inv = a.eps_av * ci / h;
                    const double si = fmax(-dv, 0.0);
                    put<T>(s, F_ALPHA, i, (alpha + dt * tau_inv * a.amin + si * dt * a.amax) / (1.0 + dt * tau_inv + si * dt));
                }
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
        if(a.stats && iters) stat_add(a.counters, grp, C_NEWTON_ITERS, iters);
        if(fails) atomicAdd((unsigned long long *)&a.counters[C_NEWTON_FAIL], (unsigned long long)fails);
        if(trunc) atomicAdd((unsigned long long *)&a.counters[C_TRUNC], (unsigned long long)trunc);
        if(a.stats && pairs) stat_add(a.counters, grp, C_GATHER_PAIRS, pairs);
    }
    __syncthreads();
    if(threadIdx.x == 0) {
        double m = sh[0];
        for(int ww = 1; ww < BLOCK / WAVE; ++ww) m = sh[ww] < m ? sh[ww] : m;
        a.partial[blockIdx.x] = m;
    }
}

static PreArgs make_args(Ctx &c)
{
    PreArgs a{};
    a.groups = c.groups;
    a.ngroups = c.ngroups;
    a.cand = c.tile_cand;
    a.mf = c.tile_mf;
    a.gnc = c.tile_nc;
    a.wcap = c.tile_wcap;
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
    a.stats = c.stats_on ? 1 : 0;
    a.partial = c.partial;
    a.prec = c.pair_rec;
    a.pm = c.posrec;  // free during the pre-interaction (the force pass repacks it)
    return a;
}

// Fields of j read by the pair sums, mirrored to FP32 before the stage
// (alpha and the Balsara switch too: the force reads their mirrors, and
// the stage rewrites them only when their switches are on).
static const int kPreInputs[] = {F_V, F_V + 1, F_V + 2, F_MASS, F_ENE, F_SOUND, F_ALPHA, F_BAL};
constexpr int kNumPreInputs = 8;

template <int D, int K, class T> static void launch_initial(Ctx &c, const PreArgs &a)
{
    hipLaunchKernelGGL((pre_initial_kernel<D, K, T>), dim3(group_grid(c)), dim3(BLOCK), 0, c.stream, c.s, make_box(c), a);
}

template <class T> static void initial_t(Ctx &c, const PreArgs &a)
{
    const int K = c.p.kernel;
    switch(c.D) {
    case 1: K ? launch_initial<1, 1, T>(c, a) : launch_initial<1, 0, T>(c, a); break;
    case 2: K ? launch_initial<2, 1, T>(c, a) : launch_initial<2, 0, T>(c, a); break;
    default: K ? launch_initial<3, 1, T>(c, a) : launch_initial<3, 0, T>(c, a); break;
    }
}

int pre_initial(Ctx &c)
{
    int rc = mirror_fields(c, kPreInputs, kNumPreInputs);
    if(rc) return rc;
    const PreArgs a = make_args(c);
    if(c.p.precision == SPHGPU_MIXED) initial_t<float>(c, a);
    else initial_t<double>(c, a);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

template <int D, int K, int TYPE, class T> static void launch_pre(Ctx &c, const PreArgs &a, const BoxDev &box)
{
    hipLaunchKernelGGL((pack_pre_kernel<D, T>), dim3(grid_for(c.n)), dim3(BLOCK), 0, c.stream, c.s, (int)c.n,
                       (PreRec<T> *)c.pair_rec, c.posrec);
    hipLaunchKernelGGL((pre_kernel<D, K, TYPE, T>), dim3(group_grid(c)), dim3(BLOCK), 0, c.stream, c.s, box, (int)c.n, a);
}

template <int D, int K, class T> static void launch_pre_t(Ctx &c, const PreArgs &a, const BoxDev &box)
{
    switch(c.p.sph_type) {
    case SPHGPU_SSPH: launch_pre<D, K, SPHGPU_SSPH, T>(c, a, box); break;
    case SPHGPU_DISPH: launch_pre<D, K, SPHGPU_DISPH, T>(c, a, box); break;
    default: launch_pre<D, K, SPHGPU_GSPH, T>(c, a, box); break;
    }
}

template <class T> static void pre_t(Ctx &c, const PreArgs &a, const BoxDev &box)
{
    const int K = c.p.kernel;
    switch(c.D) {
    case 1: K ? launch_pre_t<1, 1, T>(c, a, box) : launch_pre_t<1, 0, T>(c, a, box); break;
    case 2: K ? launch_pre_t<2, 1, T>(c, a, box) : launch_pre_t<2, 0, T>(c, a, box); break;
    default: K ? launch_pre_t<3, 1, T>(c, a, box) : launch_pre_t<3, 0, T>(c, a, box); break;
    }
}

int pre_compute(Ctx &c)
{
    int rc = mirror_fields(c, kPreInputs, kNumPreInputs);
    if(rc) return rc;
    PreArgs a = make_args(c);
    a.gpart = c.p.gravity_is_valid ? c.gpart : nullptr;
    const BoxDev box = make_box(c);
    if(c.p.precision == SPHGPU_MIXED) pre_t<float>(c, a, box);
    else pre_t<double>(c, a, box);
    SPHGPU_HIP(hipGetLastError());
    c.gpart_fresh = a.gpart != nullptr;
    return reduce_min_partials(c, group_grid(c), S_HVS);
}

}  // namespace sphgpu
