// FluidForce on the device, one thread per particle over its symmetric
// candidate list (pairs with r < max(h_i, h_j) and r != 0):
//   SSPH  src/fluid_force.cpp:26-86, viscosity :89-106, conductivity :108-116
//   DISPH src/disph/d_fluid_force.cpp:26-86 (pressure-energy form)
//   GSPH  src/gsph/g_fluid_force.cpp:39-166, limiter :27-36, HLL :168-199
// acc and dene are assigned, not accumulated, as in the reference.
#include "ctx.hpp"

namespace sphgpu {

struct ForceArgs {
    const int *list;
    const int *cnt;
    int cap;
    double gamma;
    int use_ac;
    double alpha_ac;
    int gravity;          // selects the conductivity signal speed
    int second_order;
    const double *scal;   // S_DT
    int64_t *counters;
    int *rec;             // recorded force sets (debug), same layout as list
    int *rec_cnt;
};

// van Leer limiter (src/gsph/g_fluid_force.cpp:27-36).
__device__ __forceinline__ double van_leer(double a, double b)
{
    const double ab = a * b;
    return ab <= 0.0 ? 0.0 : 2.0 * ab / (a + b);
}

// HLL with Roe-averaged wave speeds (src/gsph/g_fluid_force.cpp:168-199);
// states are (velocity along e, density, pressure, sound speed).
__device__ __forceinline__ void hll(const double L[4], const double R[4], double &pstar, double &vstar)
{
    const double sl = sqrt(L[1]), sr = sqrt(R[1]);
    const double inv = 1.0 / (sl + sr);
    const double ut = (sl * L[0] + sr * R[0]) * inv;
    const double ct = (sl * L[3] + sr * R[3]) * inv;
    const double s_l = fmin(L[0] - L[3], ut - ct);
    const double s_r = fmax(R[0] + R[3], ut + ct);
    const double c1 = L[1] * (s_l - L[0]);
    const double c2 = R[1] * (s_r - R[0]);
    const double c3 = 1.0 / (c1 - c2);
    const double c4 = L[2] - L[0] * c1;
    const double c5 = R[2] - R[0] * c2;
    vstar = (c5 - c4) * c3;
    pstar = (c1 * c5 - c2 * c4) * c3;
}

template <int D> __device__ __forceinline__ double dotd(const double a[D], const double b[D])
{
    double s = a[0] * b[0];
#pragma unroll
    for(int k = 1; k < D; ++k) s += a[k] * b[k];
    return s;
}

template <int D, int K, int TYPE>
__global__ void __launch_bounds__(BLOCK) force_kernel(State s, BoxDev box, int n, ForceArgs a)
{
    using Kern = typename KernelOf<D, K, double>::type;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const bool valid = i < n;
    int64_t pairs = 0;
    if(valid) {
        const size_t base = ((size_t)i / WAVE) * a.cap * WAVE + i % WAVE;
        const int cnt = a.cnt[i];
        double xi[D], vi[D];
#pragma unroll
        for(int k = 0; k < D; ++k) { xi[k] = s.f[F_X + k][i]; vi[k] = s.f[F_V + k][i]; }
        const double hi = s.f[F_SML][i];
        const Kern ki(hi, s.f[F_HINV][i]);
        const double mi = s.f[F_MASS][i];
        const double rhoi = s.f[F_DENS][i];
        const double pi_ = s.f[F_PRES][i];
        const double ui = s.f[F_ENE][i];
        const double ci = s.f[F_SOUND][i];
        const double alphai = s.f[F_ALPHA][i];
        const double bali = s.f[F_BAL][i];
        const double gradhi = TYPE == SPHGPU_GSPH ? 0.0 : s.f[F_GRADH][i];
        const double gm1 = a.gamma - 1.0;
        // per-formulation constants of particle i
        const double p_rho2_i = pi_ / (rhoi * rhoi);
        const double g2u_i = gm1 * gm1 * ui;
        const double g2u_p_i = g2u_i / pi_;
        const double mu_inv_i = 1.0 / (mi * ui);
        const double rho2_inv_i = 1.0 / (rhoi * rhoi);
        const double dt = a.scal[S_DT];

        double acc[D], dene = 0.0;
#pragma unroll
        for(int k = 0; k < D; ++k) acc[k] = 0.0;
        int used = 0;
        for(int k = 0; k < cnt; ++k) {
            const int j = a.list[base + (size_t)k * WAVE];
            double xj[D], rij[D];
#pragma unroll
            for(int l = 0; l < D; ++l) xj[l] = s.f[F_X + l][j];
            const double r = sqrt(sep2<D>(box, xi, xj, rij));
            const double hj = s.f[F_SML][j];
            if(r >= (hi < hj ? hj : hi) || r == 0.0) continue;
            if(a.rec) a.rec[base + (size_t)used * WAVE] = j;
            ++used;
            const Kern kj(hj, s.f[F_HINV][j]);
            const double cwi = ki.dw_coef(ki.q_of(r), r);
            const double cwj = kj.dw_coef(kj.q_of(r), r);
            double dwi[D], dwj[D], vj[D];
#pragma unroll
            for(int l = 0; l < D; ++l) {
                dwi[l] = rij[l] * cwi;
                dwj[l] = rij[l] * cwj;
                vj[l] = s.f[F_V + l][j];
            }
            const double mj = s.f[F_MASS][j];
            const double rhoj = s.f[F_DENS][j];
            const double pj = s.f[F_PRES][j];
            const double cj = s.f[F_SOUND][j];

            if constexpr(TYPE == SPHGPU_GSPH) {
                const double rinv = 1.0 / r;
                double e[D];
#pragma unroll
                for(int l = 0; l < D; ++l) e[l] = rij[l] * rinv;
                const double ve_i = dotd<D>(vi, e);
                const double ve_j = dotd<D>(vj, e);
                double pstar, vstar;
                if(a.second_order) {
                    // MUSCL reconstruction (src/gsph/g_fluid_force.cpp:98-135)
                    double R[4], L[4];
                    const double di = 0.5 * (1.0 - ci * dt * rinv);
                    const double dj = 0.5 * (1.0 - cj * dt * rinv);
                    const double dv = ve_i - ve_j;
                    double gvi[D], gvj[D];
#pragma unroll
                    for(int c = 0; c < D; ++c) {
                        double si = s.f[F_GV + 3 * c][i] * e[0], sj = s.f[F_GV + 3 * c][j] * e[0];
#pragma unroll
                        for(int l = 1; l < D; ++l) {
                            si += s.f[F_GV + 3 * c + l][i] * e[l];
                            sj += s.f[F_GV + 3 * c + l][j] * e[l];
                        }
                        gvi[c] = si;
                        gvj[c] = sj;
                    }
                    const double dve_i = dotd<D>(gvi, e) * r;
                    const double dve_j = dotd<D>(gvj, e) * r;
                    R[0] = ve_i - van_leer(dv, dve_i) * di;
                    L[0] = ve_j + van_leer(dv, dve_j) * dj;
                    double gdi = s.f[F_GD][i] * e[0], gdj = s.f[F_GD][j] * e[0];
                    double gpi = s.f[F_GP][i] * e[0], gpj = s.f[F_GP][j] * e[0];
#pragma unroll
                    for(int l = 1; l < D; ++l) {
                        gdi += s.f[F_GD + l][i] * e[l];
                        gdj += s.f[F_GD + l][j] * e[l];
                        gpi += s.f[F_GP + l][i] * e[l];
                        gpj += s.f[F_GP + l][j] * e[l];
                    }
                    const double dd = rhoi - rhoj;
                    R[1] = rhoi - van_leer(dd, gdi * r) * di;
                    L[1] = rhoj + van_leer(dd, gdj * r) * dj;
                    const double dp = pi_ - pj;
                    R[2] = pi_ - van_leer(dp, gpi * r) * di;
                    L[2] = pj + van_leer(dp, gpj * r) * dj;
                    R[3] = sqrt(a.gamma * R[2] / R[1]);
                    L[3] = sqrt(a.gamma * L[2] / L[1]);
                    hll(L, R, pstar, vstar);
                } else {
                    const double R[4] = {ve_i, rhoi, pi_, ci};
                    const double L[4] = {ve_j, rhoj, pj, cj};
                    hll(L, R, pstar, vstar);
                }
                const double fi = mj * pstar * rho2_inv_i;
                const double fj = mj * pstar / (rhoj * rhoj);
                double fdv = 0.0;
#pragma unroll
                for(int l = 0; l < D; ++l) {
                    const double f = dwi[l] * fi + dwj[l] * fj;
                    acc[l] -= f;
                    fdv += f * (e[l] * vstar - vi[l]);
                }
                dene -= fdv;
            } else {
                double vij[D], dwij[D];
#pragma unroll
                for(int l = 0; l < D; ++l) {
                    vij[l] = vi[l] - vj[l];
                    dwij[l] = (dwi[l] + dwj[l]) * 0.5;
                }
                // Monaghan (1997) viscosity (src/fluid_force.cpp:89-106)
                const double vr = dotd<D>(vij, rij);
                double pij = 0.0;
                if(vr < 0.0) {
                    const double alpha = 0.5 * (alphai + s.f[F_ALPHA][j]);
                    const double bal = 0.5 * (bali + s.f[F_BAL][j]);
                    const double w = vr / r;
                    const double vsig = ci + cj - 3.0 * w;
                    pij = -0.5 * bal * alpha * vsig * w * (2.0 / (rhoi + rhoj));
                }
                double ac = 0.0;
                if(a.use_ac) {
                    const double uj = s.f[F_ENE][j];
                    const double vs = a.gravity ? fabs(vr / r) : sqrt(2.0 * fabs(pi_ - pj) / (rhoi + rhoj));
                    ac = a.alpha_ac * mj * vs * (ui - uj) * dotd<D>(dwij, rij) / r;
                }
                const double vdwi = dotd<D>(vij, dwi);
                const double vdwij = dotd<D>(vij, dwij);
                double ci_f, cj_f, de;
                if constexpr(TYPE == SPHGPU_DISPH) {
                    const double uj = s.f[F_ENE][j];
                    const double fij = 1.0 - gradhi / (mj * uj);
                    const double fji = 1.0 - s.f[F_GRADH][j] * mu_inv_i;
                    const double u_p_j = uj / pj;
                    ci_f = mj * (g2u_p_i * uj * fij + 0.5 * pij);
                    cj_f = mj * (g2u_i * u_p_j * fji + 0.5 * pij);
                    de = mj * g2u_p_i * uj * fij * vdwi;
                } else {
                    ci_f = mj * (p_rho2_i * gradhi + 0.5 * pij);
                    cj_f = mj * (pj / (rhoj * rhoj) * s.f[F_GRADH][j] + 0.5 * pij);
                    de = mj * p_rho2_i * gradhi * vdwi;
                }
#pragma unroll
                for(int l = 0; l < D; ++l) acc[l] -= dwi[l] * ci_f + dwj[l] * cj_f;
                dene += de + 0.5 * mj * pij * vdwij + ac;
            }
        }
#pragma unroll
        for(int k = 0; k < D; ++k) s.f[F_A + k][i] = acc[k];
        s.f[F_DENE][i] = dene;
        if(a.rec_cnt) a.rec_cnt[i] = used;
        pairs = used;
    }
    pairs = wave_sum(pairs);
    if(wave_lane() == 0 && pairs) atomicAdd((unsigned long long *)&a.counters[C_FORCE_PAIRS], (unsigned long long)pairs);
}

template <int D, int K, int TYPE> static void launch_force(Ctx &c, const ForceArgs &a, const BoxDev &box)
{
    hipLaunchKernelGGL((force_kernel<D, K, TYPE>), dim3(grid_for(c.n)), dim3(BLOCK), 0, c.stream, c.s, box, (int)c.n, a);
}

template <int D, int K> static void launch_force_t(Ctx &c, const ForceArgs &a, const BoxDev &box)
{
    switch(c.p.sph_type) {
    case SPHGPU_SSPH: launch_force<D, K, SPHGPU_SSPH>(c, a, box); break;
    case SPHGPU_DISPH: launch_force<D, K, SPHGPU_DISPH>(c, a, box); break;
    default: launch_force<D, K, SPHGPU_GSPH>(c, a, box); break;
    }
}

int force_compute(Ctx &c)
{
    ForceArgs a{};
    a.list = c.list;
    a.cnt = c.list_cnt;
    a.cap = c.cap;
    a.gamma = c.p.gamma;
    a.use_ac = c.p.ac_is_valid;
    a.alpha_ac = c.p.ac_alpha;
    a.gravity = c.p.gravity_is_valid;
    a.second_order = c.p.gsph_is_2nd_order;
    a.scal = c.scal;
    a.counters = c.counters;
    a.rec = c.record ? c.rec_force : nullptr;
    a.rec_cnt = c.record ? c.rec_force_cnt : nullptr;
    const BoxDev box = make_box(c);
    const int K = c.p.kernel;
    switch(c.D) {
    case 1: K ? launch_force_t<1, 1>(c, a, box) : launch_force_t<1, 0>(c, a, box); break;
    case 2: K ? launch_force_t<2, 1>(c, a, box) : launch_force_t<2, 0>(c, a, box); break;
    default: K ? launch_force_t<3, 1>(c, a, box) : launch_force_t<3, 0>(c, a, box); break;
    }
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

}  // namespace sphgpu
