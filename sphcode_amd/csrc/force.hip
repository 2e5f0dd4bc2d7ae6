// FluidForce on the device, one wave per search group of at most 64
// particles, one lane per particle, over the group's candidate tiles: the
// tile's 64 neighbour records are staged once in LDS, and each lane then
// takes the members of its symmetric candidate set (front | back hit
// masks) among them, keeping the pairs with r < max(h_i, h_j), r != 0:
//   SSPH  src/fluid_force.cpp:26-86, viscosity :89-106, conductivity :108-116
//   DISPH src/disph/d_fluid_force.cpp:26-86 (pressure-energy form)
//   GSPH  src/gsph/g_fluid_force.cpp:39-166, limiter :27-36, HLL :168-199
// acc and dene are assigned, not accumulated, as in the reference.
//
// T is the pair arithmetic (double, or float in mixed mode: j fields from
// FP32 mirrors, separations taken in FP64 then rounded, FP32 terms
// accumulated in FP64; the pair cut is tested on the FP64 squared distance).
#include "tiles.hpp"

namespace sphgpu {

struct ForceArgs {
    const int2 *groups;
    int ngroups;
    const int *cand;        // candidate tiles (Ctx::tile_cand)
    const uint64_t *ma;     // hit masks per (tile, lane): every symmetric pair
    const int *gnc;
    int wcap;
    int cap;
    double gamma;
    int use_ac;
    double alpha_ac;
    int gravity;          // selects the conductivity signal speed
    int second_order;
    const double *scal;   // S_DT
    int64_t *counters;
    int stats;            // add to the statistics counters (errors are always counted)
    int *rec;             // recorded force sets (debug), list_at layout
    int *rec_cnt;
    const double4 *posrec;  // x, y, z, h (FP64)
    const void *frec;       // ForceRec<T>
    const void *grec;       // GradRec<T> (GSPH second order)
};

// One pass over the particles: the records the pair loop reads for j, with
// the fields the formulation reads (FP32 mirrors in mixed mode).
template <int D, int TYPE, class T>
__global__ void __launch_bounds__(BLOCK) pack_force_kernel(State s, int n, int use_u, double4 *pr, ForceRec<T> *fr,
                                                           GradRec<T> *gr)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
    double4 p;
    p.x = s.f[F_X][i];
    p.y = D > 1 ? s.f[F_X + 1][i] : 0.0;
    p.z = D > 2 ? s.f[F_X + 2][i] : 0.0;
    p.w = s.f[F_SML][i];
    pr[i] = p;
    ForceRec<T> r{};
    r.hinv = fld<T>(s, F_HINV, i);
    r.m = fld<T>(s, F_MASS, i);
    r.rho = fld<T>(s, F_DENS, i);
    r.p = fld<T>(s, F_PRES, i);
    r.c = fld<T>(s, F_SOUND, i);
#pragma unroll
    for(int l = 0; l < D; ++l) r.v[l] = fld<T>(s, F_V + l, i);
    if constexpr(TYPE != SPHGPU_GSPH) {
        r.alpha = fld<T>(s, F_ALPHA, i);
        r.bal = fld<T>(s, F_BAL, i);
        r.gradh = fld<T>(s, F_GRADH, i);
    }
    if(use_u) r.u = fld<T>(s, F_ENE, i);
    fr[i] = r;
    if constexpr(TYPE == SPHGPU_GSPH) {
        if(gr) {
            GradRec<T> g{};
#pragma unroll
            for(int l = 0; l < D; ++l) {
                g.gd[l] = fld<T>(s, F_GD + l, i);
                g.gp[l] = fld<T>(s, F_GP + l, i);
#pragma unroll
                for(int c = 0; c < D; ++c) g.gv[3 * c + l] = fld<T>(s, F_GV + 3 * c + l, i);
            }
            gr[i] = g;
        }
    }
}

// van Leer limiter (src/gsph/g_fluid_force.cpp:27-36).
template <class T> __device__ __forceinline__ T van_leer(T a, T b)
{
    const T ab = a * b;
    return ab <= T(0) ? T(0) : T(2) * ab / (a + b);
}

// HLL with Roe-averaged wave speeds (src/gsph/g_fluid_force.cpp:168-199);
// states are (velocity along e, density, pressure, sound speed).
template <class T> __device__ __forceinline__ void hll(const T L[4], const T R[4], T &pstar, T &vstar)
{
    const T sl = sqrt(L[1]), sr = sqrt(R[1]);
    const T inv = T(1) / (sl + sr);
    const T ut = (sl * L[0] + sr * R[0]) * inv;
    const T ct = (sl * L[3] + sr * R[3]) * inv;
    const T s_l = fmin(L[0] - L[3], ut - ct);
    const T s_r = fmax(R[0] + R[3], ut + ct);
    const T c1 = L[1] * (s_l - L[0]);
    const T c2 = R[1] * (s_r - R[0]);
    const T c3 = T(1) / (c1 - c2);
    const T c4 = L[2] - L[0] * c1;
    const T c5 = R[2] - R[0] * c2;
    vstar = (c5 - c4) * c3;
    pstar = (c1 * c5 - c2 * c4) * c3;
}

template <int D, class T> __device__ __forceinline__ T dotd(const T a[D], const T b[D])
{
    T s = a[0] * b[0];
#pragma unroll
    for(int k = 1; k < D; ++k) s += a[k] * b[k];
    return s;
}

template <int D, int K, int TYPE, class T>
__global__ void __launch_bounds__(BLOCK) force_kernel(State s, BoxDev box, int n, ForceArgs a)
{
    using Kern = typename KernelOf<D, K, T>::type;
    constexpr bool FP64 = sizeof(T) == 8;
    constexpr bool GS = TYPE == SPHGPU_GSPH;
    constexpr int SW = BLOCK / WAVE;
    __shared__ double4 l_p[SW][WAVE];
    __shared__ ForceRec<T> l_f[SW][WAVE];
    __shared__ GradRec<T> l_g[GS ? SW : 1][GS ? WAVE : 1];
    __shared__ int l_j[SW][WAVE];
    const int w = threadIdx.x / WAVE;
    const int lane = wave_lane();
    const int grp = xcd_block() * SW + w;
    if(grp >= a.ngroups) return;
    const int2 gr = a.groups[grp];
    const int i = gr.x + lane;
    const bool valid = i < gr.y;
    int64_t pairs = 0;
    {
        double xi[D];
        T vi[D];
        const int ii = valid ? i : gr.x;
#pragma unroll
        for(int k = 0; k < D; ++k) { xi[k] = s.f[F_X + k][ii]; vi[k] = (T)s.f[F_V + k][ii]; }
        const double hi_d = s.f[F_SML][ii];
        const T hi = (T)hi_d;
        const Kern ki(hi, (T)s.f[F_HINV][ii]);
        const T mi = (T)s.f[F_MASS][ii];
        const T rhoi = (T)s.f[F_DENS][ii];
        const T pi_ = (T)s.f[F_PRES][ii];
        const T ui = (T)s.f[F_ENE][ii];
        const T ci = (T)s.f[F_SOUND][ii];
        const T alphai = (T)s.f[F_ALPHA][ii];
        const T bali = (T)s.f[F_BAL][ii];
        const T gradhi = TYPE == SPHGPU_GSPH ? T(0) : (T)s.f[F_GRADH][ii];
        const T gm1 = (T)(a.gamma - 1.0);
        const T gam = (T)a.gamma;
        // per-formulation constants of particle i
        const T p_rho2_i = pi_ / (rhoi * rhoi);
        const T g2u_i = gm1 * gm1 * ui;
        const T g2u_p_i = g2u_i / pi_;
        const T mu_inv_i = T(1) / (mi * ui);
        const T rho2_inv_i = T(1) / (rhoi * rhoi);
        const T dt = (T)a.scal[S_DT];
        GradRec<T> gri{};
        if constexpr(GS)
            if(a.second_order) gri = ((const GradRec<T> *)a.grec)[ii];

        Acc<T> acc[D], dene;
        int used = 0;
        const ForceRec<T> *frec = (const ForceRec<T> *)a.frec;
        const size_t gbase = (size_t)grp * a.wcap;
        const int nc = a.gnc[grp];
        const int ntile = (nc + WAVE - 1) / WAVE;
        auto stage = [&](int tt) -> uint64_t {
            const size_t o = gbase + (size_t)tt * WAVE;
            // lane k loads candidate k's records
            const int jl = stream_ld(a.cand + o + lane);
            if(jl >= 0) {
                l_p[w][lane] = a.posrec[jl];
                l_f[w][lane] = frec[jl];
                if constexpr(GS)
                    if(a.second_order) l_g[GS ? w : 0][lane] = ((const GradRec<T> *)a.grec)[jl];
            }
            l_j[w][lane] = jl;
            return valid ? stream_ld(a.ma + o + lane) : 0ull;
        };
        auto pair = [&](int kc) {
            const int j = l_j[w][kc];
            const double4 pj4 = l_p[w][kc];
            double xj[D], d[D];
            xj[0] = pj4.x;
            if constexpr(D > 1) xj[1] = pj4.y;
            if constexpr(D > 2) xj[2] = pj4.z;
            const double r2 = sep2<D>(box, xi, xj, d);
            const double hj_d = pj4.w;
            const double hmax = hi_d < hj_d ? hj_d : hi_d;
            T r;
            if constexpr(FP64) {
                r = sqrt(r2);
                if(r >= hmax || r == 0.0) return;  // src/fluid_force.cpp:62-64
            } else {
                if(!(r2 < hmax * hmax) || r2 == 0.0) return;
                r = sqrtf((float)r2);
            }
            if(a.rec && used < a.cap) a.rec[list_at(i, used, a.cap)] = j;
            ++used;
            T rij[D];
#pragma unroll
            for(int l = 0; l < D; ++l) rij[l] = (T)d[l];
            const ForceRec<T> fj = l_f[w][kc];
            const Kern kj(T(0), fj.hinv);  // the kernel object reads 1/h only
            const T cwi = ki.dw_coef(ki.q_of(r), r);
            const T cwj = kj.dw_coef(kj.q_of(r), r);
            T dwi[D], dwj[D], vj[D];
#pragma unroll
            for(int l = 0; l < D; ++l) {
                dwi[l] = rij[l] * cwi;
                dwj[l] = rij[l] * cwj;
                vj[l] = fj.v[l];
            }
            const T mj = fj.m;
            const T rhoj = fj.rho;
            const T pj = fj.p;
            const T cj = fj.c;

            if constexpr(TYPE == SPHGPU_GSPH) {
                const T rinv = T(1) / r;
                T e[D];
#pragma unroll
                for(int l = 0; l < D; ++l) e[l] = rij[l] * rinv;
                const T ve_i = dotd<D, T>(vi, e);
                const T ve_j = dotd<D, T>(vj, e);
                T pstar, vstar;
                if(a.second_order) {
                    // MUSCL reconstruction (src/gsph/g_fluid_force.cpp:98-135)
                    T R[4], L[4];
                    const T di = T(0.5) * (T(1) - ci * dt * rinv);
                    const T dj = T(0.5) * (T(1) - cj * dt * rinv);
                    const T dv = ve_i - ve_j;
                    const GradRec<T> &grj = l_g[GS ? w : 0][GS ? kc : 0];
                    T gvi[D], gvj[D];
#pragma unroll
                    for(int c = 0; c < D; ++c) {
                        T si = gri.gv[3 * c] * e[0], sj = grj.gv[3 * c] * e[0];
#pragma unroll
                        for(int l = 1; l < D; ++l) {
                            si += gri.gv[3 * c + l] * e[l];
                            sj += grj.gv[3 * c + l] * e[l];
                        }
                        gvi[c] = si;
                        gvj[c] = sj;
                    }
                    const T dve_i = dotd<D, T>(gvi, e) * r;
                    const T dve_j = dotd<D, T>(gvj, e) * r;
                    R[0] = ve_i - van_leer(dv, dve_i) * di;
                    L[0] = ve_j + van_leer(dv, dve_j) * dj;
                    T gdi = gri.gd[0] * e[0], gdj = grj.gd[0] * e[0];
                    T gpi = gri.gp[0] * e[0], gpj = grj.gp[0] * e[0];
#pragma unroll
                    for(int l = 1; l < D; ++l) {
                        gdi += gri.gd[l] * e[l];
                        gdj += grj.gd[l] * e[l];
                        gpi += gri.gp[l] * e[l];
                        gpj += grj.gp[l] * e[l];
                    }
                    const T dd = rhoi - rhoj;
                    R[1] = rhoi - van_leer(dd, gdi * r) * di;
                    L[1] = rhoj + van_leer(dd, gdj * r) * dj;
                    const T dp = pi_ - pj;
                    R[2] = pi_ - van_leer(dp, gpi * r) * di;
                    L[2] = pj + van_leer(dp, gpj * r) * dj;
                    R[3] = sqrt(gam * R[2] / R[1]);
                    L[3] = sqrt(gam * L[2] / L[1]);
                    hll(L, R, pstar, vstar);
                } else {
                    const T R[4] = {ve_i, rhoi, pi_, ci};
                    const T L[4] = {ve_j, rhoj, pj, cj};
                    hll(L, R, pstar, vstar);
                }
                const T fi = mj * pstar * rho2_inv_i;
                const T fj = mj * pstar / (rhoj * rhoj);
                T fdv = T(0);
#pragma unroll
                for(int l = 0; l < D; ++l) {
                    const T f = dwi[l] * fi + dwj[l] * fj;
                    acc[l].add(-f);
                    fdv += f * (e[l] * vstar - vi[l]);
                }
                dene.add(-fdv);
            } else {
                T vij[D], dwij[D];
#pragma unroll
                for(int l = 0; l < D; ++l) {
                    vij[l] = vi[l] - vj[l];
                    dwij[l] = (dwi[l] + dwj[l]) * T(0.5);
                }
                // Monaghan (1997) viscosity (src/fluid_force.cpp:89-106)
                const T vr = dotd<D, T>(vij, rij);
                T pij = T(0);
                if(vr < T(0)) {
                    const T alpha = T(0.5) * (alphai + fj.alpha);
                    const T bal = T(0.5) * (bali + fj.bal);
                    const T w = vr / r;
                    const T vsig = ci + cj - T(3) * w;
                    pij = T(-0.5) * bal * alpha * vsig * w * (T(2) / (rhoi + rhoj));
                }
                T ac = T(0);
                if(a.use_ac) {
                    const T uj = fj.u;
                    const T vs = a.gravity ? fabs(vr / r) : sqrt(T(2) * fabs(pi_ - pj) / (rhoi + rhoj));
                    ac = (T)a.alpha_ac * mj * vs * (ui - uj) * dotd<D, T>(dwij, rij) / r;
                }
                const T vdwi = dotd<D, T>(vij, dwi);
                const T vdwij = dotd<D, T>(vij, dwij);
                T ci_f, cj_f, de;
                if constexpr(TYPE == SPHGPU_DISPH) {
                    const T uj = fj.u;
                    const T fij = T(1) - gradhi / (mj * uj);
                    const T fji = T(1) - fj.gradh * mu_inv_i;
                    const T u_p_j = uj / pj;
                    ci_f = mj * (g2u_p_i * uj * fij + T(0.5) * pij);
                    cj_f = mj * (g2u_i * u_p_j * fji + T(0.5) * pij);
                    de = mj * g2u_p_i * uj * fij * vdwi;
                } else {
                    ci_f = mj * (p_rho2_i * gradhi + T(0.5) * pij);
                    cj_f = mj * (pj / (rhoj * rhoj) * fj.gradh + T(0.5) * pij);
                    de = mj * p_rho2_i * gradhi * vdwi;
                }
#pragma unroll
                for(int l = 0; l < D; ++l) acc[l].add(-(dwi[l] * ci_f + dwj[l] * cj_f));
                dene.add(de + T(0.5) * mj * pij * vdwij + ac);
            }
        };
        // the cut-off r < max(h_i, h_j), r > 0 of the pair, tested ahead
        auto keep = [&](int kc) -> bool {
            const double4 pj4 = l_p[w][kc];
            double xj[D], d[D];
            xj[0] = pj4.x;
            if constexpr(D > 1) xj[1] = pj4.y;
            if constexpr(D > 2) xj[2] = pj4.z;
            const double r2 = sep2<D>(box, xi, xj, d);
            const double hmax = hi_d < pj4.w ? pj4.w : hi_d;
            if constexpr(FP64) {
                const double r = sqrt(r2);
                return !(r >= hmax || r == 0.0);
            } else {
                return r2 < hmax * hmax && r2 != 0.0;
            }
        };
        // refined in 3D only, see tiles.hpp
        if constexpr(D == 3) {
            auto refine = [&](uint64_t m) { return refine_bits(m, keep); };
            tile_sweep(ntile, stage, refine, pair);
        } else {
            tile_sweep(ntile, stage, pair);
        }
        if(valid) {
#pragma unroll
            for(int k = 0; k < D; ++k) s.f[F_A + k][i] = acc[k].val();
            s.f[F_DENE][i] = dene.val();
            if(a.rec_cnt) a.rec_cnt[i] = used < a.cap ? used : a.cap;
            pairs = used;
        }
    }
    if(a.stats) {
        pairs = wave_sum(pairs);
        if(wave_lane() == 0 && pairs) stat_add(a.counters, grp, C_FORCE_PAIRS, pairs);
    }
}

template <int D, int K, int TYPE, class T> static void launch_force(Ctx &c, const ForceArgs &a, const BoxDev &box)
{
    const int use_u = TYPE == SPHGPU_DISPH || a.use_ac;
    hipLaunchKernelGGL((pack_force_kernel<D, TYPE, T>), dim3(grid_for(c.n)), dim3(BLOCK), 0, c.stream, c.s, (int)c.n, use_u,
                       c.posrec, (ForceRec<T> *)c.pair_rec, (GradRec<T> *)c.grad_rec);
    if(c.after_pack) c.after_pack_rc = c.after_pack(c);
    hipLaunchKernelGGL((force_kernel<D, K, TYPE, T>), dim3(group_grid(c)), dim3(BLOCK), 0, c.stream, c.s, box, (int)c.n, a);
}

template <int D, int K, class T> static void launch_force_t(Ctx &c, const ForceArgs &a, const BoxDev &box)
{
    switch(c.p.sph_type) {
    case SPHGPU_SSPH: launch_force<D, K, SPHGPU_SSPH, T>(c, a, box); break;
    case SPHGPU_DISPH: launch_force<D, K, SPHGPU_DISPH, T>(c, a, box); break;
    default: launch_force<D, K, SPHGPU_GSPH, T>(c, a, box); break;
    }
}

template <class T> static void force_t(Ctx &c, const ForceArgs &a, const BoxDev &box)
{
    const int K = c.p.kernel;
    switch(c.D) {
    case 1: K ? launch_force_t<1, 1, T>(c, a, box) : launch_force_t<1, 0, T>(c, a, box); break;
    case 2: K ? launch_force_t<2, 1, T>(c, a, box) : launch_force_t<2, 0, T>(c, a, box); break;
    default: K ? launch_force_t<3, 1, T>(c, a, box) : launch_force_t<3, 0, T>(c, a, box); break;
    }
}

int force_compute(Ctx &c)
{
    ForceArgs a{};
    a.groups = c.groups;
    a.ngroups = c.ngroups;
    a.cand = c.tile_cand;
    a.ma = c.tile_ma;
    a.gnc = c.tile_nc;
    a.wcap = c.tile_wcap;
    a.cap = c.cap;
    a.gamma = c.p.gamma;
    a.use_ac = c.p.ac_is_valid;
    a.alpha_ac = c.p.ac_alpha;
    a.gravity = c.p.gravity_is_valid;
    a.second_order = c.p.gsph_is_2nd_order;
    a.scal = c.scal;
    a.counters = c.counters;
    a.stats = c.stats_on ? 1 : 0;
    a.rec = c.record ? c.rec_force : nullptr;
    a.rec_cnt = c.record ? c.rec_force_cnt : nullptr;
    a.posrec = c.posrec;
    a.frec = c.pair_rec;
    a.grec = c.grad_rec;
    const BoxDev box = make_box(c);
    c.after_pack_rc = 0;
    if(c.p.precision == SPHGPU_MIXED) force_t<float>(c, a, box);
    else force_t<double>(c, a, box);
    SPHGPU_HIP(hipGetLastError());
    return c.after_pack_rc;
}

}  // namespace sphgpu
