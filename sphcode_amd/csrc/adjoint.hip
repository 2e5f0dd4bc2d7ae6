// Adjoint of the SSPH step for the differentiable rollout (a capability the
// reference does not have).  The forward map of one step with a fixed dt is
//
//   v_p = v + a dt/2,  u_p = u + du dt/2                 (src/solver.cpp:431-456)
//   x'  = x + v_p dt,  v~ = v + a dt,  u~ = u + du dt,  c~ = sqrt(g(g-1) u~)
//   (a', du') = F(x', v~, u~)      PreInteraction + FluidForce
//                                  (src/pre_interaction.cpp:39-169, src/fluid_force.cpp:26-116)
//   v'  = v_p + a' dt/2,  u' = u_p + du' dt/2           (src/solver.cpp:458-474)
//
// and Solver::initialize evaluates F once at the initial state.  F is
// differentiated with the neighbour lists of the forward evaluation held
// fixed (the kernels vanish with their first derivatives at the support
// edge, so F is continuous across list changes), and h enters through the
// implicit function theorem on G(h, x) = rho(h, x) h^D - m N_ngb / A_D = 0,
// which the Newton iteration solves (src/pre_interaction.cpp:227-283).
//
// All pair sums of the adjoint are gathers over the symmetric force list
// (r < max(h_i, h_j)), which contains every pair of the gather sets
// (r < h_i) in either direction: particle k collects the derivative of every
// pair term with respect to its own variables, so no atomics are needed.
// Three passes per force evaluation:
//   1. force pairs  -> adjoints of x, v (direct), h, rho, A = P Omega / rho^2,
//                      Balsara switch, sound speed
//   2. gather sums  -> chain through P, Omega, Balsara, h (IFT) to seeds on
//                      the sums rho, d rho/dh, div v, curl v
//   3. gather pairs -> adjoints of x and v through those sums
// The pair arithmetic follows the context's precision (FP64, or FP32 with
// FP64 accumulation); the per-particle chain rule and the IFT are FP64.
#include <vector>

#include "ctx.hpp"
#include "kderiv.hpp"

namespace sphgpu {

// work arrays (device order), W(k, i) = work[k * n + i]
enum Work : int {
    W_AF = 0,      // seed on a (3)
    W_DUF = 3,     // seed on du
    W_XB = 4,      // adjoint of x (3)
    W_VB = 7,      // adjoint of v (3)
    W_HB = 10,     // partial adjoint of h
    W_RHOB = 11,   // partial adjoint of rho
    W_AB = 12,     // adjoint of A = P Omega / rho^2
    W_BB = 13,     // adjoint of the Balsara switch
    W_CB = 14,     // adjoint of the sound speed
    W_SRHO = 15,   // seed on sum m W
    W_SRHOP = 16,  // seed on sum m dW/dh
    W_SDV = 17,    // seed on sum m v_ij . grad W
    W_SRV = 18,    // seed on sum m v_ij x grad W (3)
    W_UB = 21,     // adjoint of u
    NWORK = 22
};

// checkpointed fields: what one step or Solver::initialize reads
static int ck_fields(int D, int *f)
{
    int k = 0;
    for(int a = 0; a < D; ++a) f[k++] = F_X + a;
    for(int a = 0; a < D; ++a) f[k++] = F_V + a;
    for(int a = 0; a < D; ++a) f[k++] = F_A + a;
    const int s[] = {F_MASS, F_ENE, F_DENE, F_SML, F_DENS, F_ALPHA, F_BAL, F_SOUND};
    for(int v : s) f[k++] = v;
    return k;
}

struct AdjArgs {
    const int *list;
    const int *cnt;
    const int *cnt2;
    int cap;
    double *w;       // work arrays
    int64_t n;       // their stride
    double *adj;     // adjoint state, caller order
    double gamma;
    double nngb;     // neighbor_number (for b in the IFT)
    int balsara;
    double dt;
    int kick;
};

__device__ __forceinline__ double &W(const AdjArgs &a, int k, int i) { return a.w[(int64_t)k * a.n + i]; }

// layout of the adjoint state (caller order): x (D), v (D), u, a (D), du
template <int D> __device__ __forceinline__ double &ADJ(const AdjArgs &a, int comp, int id)
{
    return a.adj[(int64_t)comp * a.n + id];
}
template <int D> struct AdjLayout {
    static constexpr int X = 0, V = D, U = 2 * D, A = 2 * D + 1, DU = 3 * D + 1, N = 3 * D + 2;
};

// seeds of F from the adjoint of the step's output state
template <int D> __global__ void __launch_bounds__(BLOCK) adj_seed_kernel(State s, int n, AdjArgs a)
{
    using L = AdjLayout<D>;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
    const int id = s.id[i];
    const double half = a.kick ? 0.5 * a.dt : 0.0;
#pragma unroll
    for(int k = 0; k < D; ++k) W(a, W_AF + k, i) = ADJ<D>(a, L::A + k, id) + half * ADJ<D>(a, L::V + k, id);
    W(a, W_DUF, i) = ADJ<D>(a, L::DU, id) + half * ADJ<D>(a, L::U, id);
}

// Pair arithmetic of the passes: T = double, or float in mixed precision
// (separations taken in FP64 and rounded, fields of j from the FP32
// mirrors, terms accumulated in FP64 by Acc<T>); the pair-set tests are
// those of the forward stages in the same precision.
template <class T> __device__ __forceinline__ T pair_r(double r2)
{
    if constexpr(sizeof(T) == 8) return sqrt(r2);
    else return gather_r_f32(r2);
}

// pass 1: force pairs (src/fluid_force.cpp:26-106)
template <int D, int K, class T>
__global__ void __launch_bounds__(BLOCK) adj_force_kernel(State s, BoxDev box, int n, AdjArgs a)
{
    constexpr bool FP64 = sizeof(T) == 8;
    const int i = xcd_block() * BLOCK + threadIdx.x;
    if(i >= n) return;
    const int cnt = a.cnt[i];
    const int total = cnt + a.cnt2[i];
    double xi[D];
    T vi[D], afi[D];
#pragma unroll
    for(int k = 0; k < D; ++k) {
        xi[k] = s.f[F_X + k][i];
        vi[k] = (T)s.f[F_V + k][i];
        afi[k] = (T)W(a, W_AF + k, i);
    }
    const double hi_d = s.f[F_SML][i];
    const T hi = (T)hi_d, rhoi = (T)s.f[F_DENS][i], pi_ = (T)s.f[F_PRES][i], omi = (T)s.f[F_GRADH][i];
    const T bi = (T)s.f[F_BAL][i], ci = (T)s.f[F_SOUND][i], ali = (T)s.f[F_ALPHA][i], mi = (T)s.f[F_MASS][i];
    const T dufi = (T)W(a, W_DUF, i);
    const T Ai = pi_ * omi / (rhoi * rhoi);

    Acc<T> xb[D], vb[D], hb, rhob, Ab, Bb, cb;
    for(int kk = 0; kk < total; ++kk) {
        const int j = a.list[list_at(i, list_slot(kk, cnt, a.cap), a.cap)];
        double xj[D], dd[D];
#pragma unroll
        for(int k = 0; k < D; ++k) xj[k] = s.f[F_X + k][j];
        const double r2 = sep2<D>(box, xi, xj, dd);
        const double hj_d = s.f[F_SML][j];
        const double hmax = hi_d < hj_d ? hj_d : hi_d;
        T r;
        if constexpr(FP64) {
            r = sqrt(r2);
            if(r >= hmax || r == 0.0) continue;
        } else {
            if(!(r2 < hmax * hmax) || r2 == 0.0) continue;
            r = sqrtf((float)r2);
        }
        T d[D];
#pragma unroll
        for(int k = 0; k < D; ++k) d[k] = (T)dd[k];
        const T hj = fld<T>(s, F_SML, j);
        const T rhoj = fld<T>(s, F_DENS, j), pj = fld<T>(s, F_PRES, j), omj = fld<T>(s, F_GRADH, j);
        const T bj = fld<T>(s, F_BAL, j), cj = fld<T>(s, F_SOUND, j), alj = fld<T>(s, F_ALPHA, j);
        const T mj = fld<T>(s, F_MASS, j);
        const T dufj = (T)W(a, W_DUF, j);
        const T Aj = pj * omj / (rhoj * rhoj);
        const KernelDerivs<T> ki = kernel_derivs<D, K, T>(r, hi);
        const KernelDerivs<T> kj = kernel_derivs<D, K, T>(r, hj);
        const T gi = ki.f, gj = kj.f, gb = T(0.5) * (gi + gj);
        T v[D], beta[D];
        T sv = T(0), b = T(0);
#pragma unroll
        for(int k = 0; k < D; ++k) {
            v[k] = vi[k] - fld<T>(s, F_V + k, j);
            beta[k] = mi * (T)W(a, W_AF + k, j) - mj * afi[k];
            sv += v[k] * d[k];
            b += beta[k] * d[k];
        }
        const T psii = mj * dufi, psij = mi * dufj, psis = psii + psij;
        // Monaghan viscosity Pi = K (C - 3 w) w for w = v.r / r < 0
        T Pi = T(0), Kv = T(0), w = T(0);
        const T Cs = ci + cj, rhos = rhoi + rhoj, al2 = ali + alj;
        if(sv < T(0)) {
            w = sv / r;
            Kv = T(-0.25) * (bi + bj) * al2 / rhos;
            Pi = Kv * (Cs - T(3) * w) * w;
        }
        const T Phi = Ai * gi + Aj * gj + Pi * gb;
        const T EPi = (b + T(0.5) * sv * psis) * gb;
        const T Egi = b * (Ai + T(0.5) * Pi) + sv * (psii * Ai + T(0.25) * psis * Pi);
        const T Egj = b * (Aj + T(0.5) * Pi) + sv * (psij * Aj + T(0.25) * psis * Pi);
        T Es = psii * Ai * gi + psij * Aj * gj + T(0.5) * psis * Pi * gb;
        T Er = Egi * ki.fr + Egj * kj.fr;
        if(sv < T(0)) {
            const T Piw = Kv * (Cs - T(6) * w);
            Es += EPi * Piw / r;
            Er -= EPi * Piw * w / r;
            rhob.add(-(EPi * Pi / rhos));
            cb.add(EPi * Kv * w);
            Bb.add(EPi * (Cs - T(3) * w) * w * (T(-0.25) * al2 / rhos));
        }
        const T er = Er / r;
#pragma unroll
        for(int k = 0; k < D; ++k) {
            xb[k].add(beta[k] * Phi + v[k] * Es + d[k] * er);
            vb[k].add(d[k] * Es);
        }
        hb.add(Egi * ki.fh);
        Ab.add((b + sv * psii) * gi);
    }
#pragma unroll
    for(int k = 0; k < D; ++k) {
        W(a, W_XB + k, i) = xb[k].val();
        W(a, W_VB + k, i) = vb[k].val();
    }
    W(a, W_HB, i) = hb.val();
    W(a, W_RHOB, i) = rhob.val();
    W(a, W_AB, i) = Ab.val();
    W(a, W_BB, i) = Bb.val();
    W(a, W_CB, i) = cb.val();
}

// v x r for D = 3 (3 components) and D = 2 (z only, stored in slot 2)
template <int D, class T> __device__ __forceinline__ void cross_vr(const T v[D], const T d[D], T o[3])
{
    o[0] = o[1] = o[2] = T(0);
    if constexpr(D == 3) {
        o[0] = v[1] * d[2] - v[2] * d[1];
        o[1] = v[2] * d[0] - v[0] * d[2];
        o[2] = v[0] * d[1] - v[1] * d[0];
    } else if constexpr(D == 2) {
        o[2] = v[0] * d[1] - v[1] * d[0];
    }
}

// pass 2: gather sums of particle i and the chain rule through its
// derived quantities (src/pre_interaction.cpp:66-169), the latter in FP64
template <int D, int K, class T>
__global__ void __launch_bounds__(BLOCK) adj_pre_kernel(State s, BoxDev box, int n, AdjArgs a)
{
    const int i = xcd_block() * BLOCK + threadIdx.x;
    if(i >= n) return;
    const int cnt = a.cnt[i];
    const int total = cnt + a.cnt2[i];
    double xi[D];
    T vi[D];
#pragma unroll
    for(int k = 0; k < D; ++k) {
        xi[k] = s.f[F_X + k][i];
        vi[k] = (T)s.f[F_V + k][i];
    }
    const double h = s.f[F_SML][i];
    const T ht = (T)h;
    Acc<T> rho_a, rhop_a, rhopp_a, Dv_a, Dvh_a, Rv_a[3], Rvh_a[3];
    for(int kk = 0; kk < total; ++kk) {
        const int j = a.list[list_at(i, list_slot(kk, cnt, a.cap), a.cap)];
        double xj[D], dd[D];
#pragma unroll
        for(int k = 0; k < D; ++k) xj[k] = s.f[F_X + k][j];
        const double r2 = sep2<D>(box, xi, xj, dd);
        const T r = pair_r<T>(r2);
        if(!(r < ht)) continue;
        const T mj = fld<T>(s, F_MASS, j);
        const KernelDerivs<T> kd = kernel_derivs<D, K, T>(r, ht);
        rho_a.add(mj * kd.w);
        rhop_a.add(mj * kd.wh);
        rhopp_a.add(mj * kd.whh);
        if(j != i && a.balsara && D > 1) {
            T v[D], d[D], sv = T(0);
#pragma unroll
            for(int k = 0; k < D; ++k) {
                d[k] = (T)dd[k];
                v[k] = vi[k] - fld<T>(s, F_V + k, j);
                sv += v[k] * d[k];
            }
            T cr[3];
            cross_vr<D, T>(v, d, cr);
            Dv_a.add(mj * sv * kd.f);
            Dvh_a.add(mj * sv * kd.fh);
#pragma unroll
            for(int k = 0; k < 3; ++k) {
                Rv_a[k].add(mj * cr[k] * kd.f);
                Rvh_a[k].add(mj * cr[k] * kd.fh);
            }
        }
    }
    const double rho = rho_a.val(), rhop = rhop_a.val(), rhopp = rhopp_a.val(), Dv = Dv_a.val(), Dvh = Dvh_a.val();
    const double Rv[3] = {Rv_a[0].val(), Rv_a[1].val(), Rv_a[2].val()};
    const double Rvh[3] = {Rvh_a[0].val(), Rvh_a[1].val(), Rvh_a[2].val()};

    const double gm1 = a.gamma - 1.0;
    const double rho_f = s.f[F_DENS][i], pres = s.f[F_PRES][i], om = s.f[F_GRADH][i];
    const double ci = s.f[F_SOUND][i], ui = s.f[F_ENE][i];
    const double Ab = W(a, W_AB, i);
    const double A = pres * om / (rho_f * rho_f);
    // A = P Omega / rho^2
    const double Pb = Ab * om / (rho_f * rho_f);
    const double Omb = Ab * pres / (rho_f * rho_f);
    double rhob = W(a, W_RHOB, i) - 2.0 * Ab * A / rho_f;
    double hb = W(a, W_HB, i);
    double cb = W(a, W_CB, i);
    // P = (gamma - 1) rho u
    double ub = gm1 * rho_f * Pb;
    rhob += gm1 * ui * Pb;
    // Omega = 1 / (1 + X), X = h rho' / (D rho)
    const double X = h * rhop / (D * rho);
    const double Xb = -om * om * Omb;
    hb += Xb * rhop / (D * rho);
    const double rhopb = Xb * h / (D * rho);
    rhob -= Xb * X / rho;
    // Balsara switch B = |div| / (|div| + |rot| + 1e-4 c / h)
    double Dvb = 0.0, Rvb[3] = {0.0, 0.0, 0.0};
    if(a.balsara && D > 1) {
        const double Bb = W(a, W_BB, i);
        const double div = -Dv / rho;
        const double R = sqrt(Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2]);
        const double rot = R / rho;
        const double ad = fabs(div);
        const double ec = 1e-4 * ci / h;
        const double den = ad + rot + ec;
        const double den2 = den * den;
        const double adb = Bb * (rot + ec) / den2;
        const double rotb = -Bb * ad / den2;
        cb -= Bb * ad * (1e-4 / h) / den2;
        hb += Bb * ad * ec / (h * den2);
        const double divb = div > 0.0 ? adb : div < 0.0 ? -adb : 0.0;
        Dvb = -divb / rho;
        rhob -= divb * div / rho;
        if(R > 0.0) {
#pragma unroll
            for(int k = 0; k < 3; ++k) Rvb[k] = rotb * Rv[k] / (R * rho);
            rhob -= rotb * rot / rho;
        }
    }
    // every h dependence of the sums, then h through G(h, x) = rho h^D - b
    hb += rhob * rhop + rhopb * rhopp + Dvb * Dvh + Rvb[0] * Rvh[0] + Rvb[1] * Rvh[1] + Rvb[2] * Rvh[2];
    const double hD1 = ipow<double, D - 1>(h);
    const double Gh = rhop * hD1 * h + D * rho * hD1;
    const double g = -hb / Gh;
    W(a, W_SRHO, i) = rhob + g * hD1 * h;
    W(a, W_SRHOP, i) = rhopb;
    W(a, W_SDV, i) = Dvb;
#pragma unroll
    for(int k = 0; k < 3; ++k) W(a, W_SRV + k, i) = Rvb[k];
    // c = sqrt(gamma (gamma - 1) u)
    if(ci > 0.0) ub += cb * a.gamma * gm1 / (2.0 * ci);
    W(a, W_UB, i) = ub;
}

// derivative of the seeded sum terms of one pair with respect to r_ij and
// v_ij; the terms are even in (r_ij, v_ij), so both ends use the same form
template <int D, int K, class T>
__device__ __forceinline__ void sum_term(T r, const T d[D], const T v[D], T sv, T m, T h, T srho, T srhop, T sdv,
                                         const T srv[3], Acc<T> xb[D], Acc<T> vb[D])
{
    const KernelDerivs<T> kd = kernel_derivs<D, K, T>(r, h);
    T gQd[D], gQv[D], Q = T(0);
    if constexpr(D == 3) {
        // Q = srv . (v x r):  dQ/dr = srv x v,  dQ/dv = r x srv
        gQd[0] = srv[1] * v[2] - srv[2] * v[1];
        gQd[1] = srv[2] * v[0] - srv[0] * v[2];
        gQd[2] = srv[0] * v[1] - srv[1] * v[0];
        gQv[0] = d[1] * srv[2] - d[2] * srv[1];
        gQv[1] = d[2] * srv[0] - d[0] * srv[2];
        gQv[2] = d[0] * srv[1] - d[1] * srv[0];
        Q = gQd[0] * d[0] + gQd[1] * d[1] + gQd[2] * d[2];
    } else if constexpr(D == 2) {
        gQd[0] = -srv[2] * v[1];
        gQd[1] = srv[2] * v[0];
        gQv[0] = srv[2] * d[1];
        gQv[1] = -srv[2] * d[0];
        Q = gQd[0] * d[0] + gQd[1] * d[1];
    } else {
        gQd[0] = gQv[0] = T(0);
    }
    const T radial = m * (srho * kd.f * r + srhop * kd.whr + (sdv * sv + Q) * kd.fr) / r;
    const T mf = m * kd.f;
#pragma unroll
    for(int k = 0; k < D; ++k) {
        xb[k].add(radial * d[k] + mf * (sdv * v[k] + gQd[k]));
        vb[k].add(mf * (sdv * d[k] + gQv[k]));
    }
}

// pass 3: adjoint of x and v through the seeded gather sums of i (pairs
// with r < h_i) and of its neighbours j (pairs with r < h_j)
template <int D, int K, class T>
__global__ void __launch_bounds__(BLOCK) adj_sum_kernel(State s, BoxDev box, int n, AdjArgs a)
{
    const int i = xcd_block() * BLOCK + threadIdx.x;
    if(i >= n) return;
    const int cnt = a.cnt[i];
    const int total = cnt + a.cnt2[i];
    double xi[D];
    T vi[D];
#pragma unroll
    for(int k = 0; k < D; ++k) {
        xi[k] = s.f[F_X + k][i];
        vi[k] = (T)s.f[F_V + k][i];
    }
    const T hi = (T)s.f[F_SML][i], mi = (T)s.f[F_MASS][i];
    const T srho_i = (T)W(a, W_SRHO, i), srhop_i = (T)W(a, W_SRHOP, i), sdv_i = (T)W(a, W_SDV, i);
    const T srv_i[3] = {(T)W(a, W_SRV, i), (T)W(a, W_SRV + 1, i), (T)W(a, W_SRV + 2, i)};
    Acc<T> xb[D], vb[D];
    for(int kk = 0; kk < total; ++kk) {
        const int j = a.list[list_at(i, list_slot(kk, cnt, a.cap), a.cap)];
        double xj[D], dd[D];
#pragma unroll
        for(int k = 0; k < D; ++k) xj[k] = s.f[F_X + k][j];
        const T r = pair_r<T>(sep2<D>(box, xi, xj, dd));
        const T hj = fld<T>(s, F_SML, j);
        const bool in_i = r < hi, in_j = r < hj;
        if(r == T(0) || !(in_i || in_j)) continue;
        T v[D], d[D], sv = T(0);
#pragma unroll
        for(int k = 0; k < D; ++k) {
            d[k] = (T)dd[k];
            v[k] = vi[k] - fld<T>(s, F_V + k, j);
            sv += v[k] * d[k];
        }
        if(in_i) sum_term<D, K, T>(r, d, v, sv, fld<T>(s, F_MASS, j), hi, srho_i, srhop_i, sdv_i, srv_i, xb, vb);
        if(in_j) {
            const T srv_j[3] = {(T)W(a, W_SRV, j), (T)W(a, W_SRV + 1, j), (T)W(a, W_SRV + 2, j)};
            sum_term<D, K, T>(r, d, v, sv, mi, hj, (T)W(a, W_SRHO, j), (T)W(a, W_SRHOP, j), (T)W(a, W_SDV, j), srv_j,
                              xb, vb);
        }
    }
#pragma unroll
    for(int k = 0; k < D; ++k) {
        W(a, W_XB + k, i) += xb[k].val();
        W(a, W_VB + k, i) += vb[k].val();
    }
}

// kick-drift adjoint and write-back in caller order
template <int D> __global__ void __launch_bounds__(BLOCK) adj_state_kernel(State s, int n, AdjArgs a)
{
    using L = AdjLayout<D>;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
    const int id = s.id[i];
    const double dt = a.dt;
    const double ubF = W(a, W_UB, i);
    if(a.kick) {
#pragma unroll
        for(int k = 0; k < D; ++k) {
            const double xbar = ADJ<D>(a, L::X + k, id) + W(a, W_XB + k, i);
            const double vtb = W(a, W_VB + k, i);
            const double vpb = ADJ<D>(a, L::V + k, id) + xbar * dt;
            ADJ<D>(a, L::X + k, id) = xbar;
            ADJ<D>(a, L::V + k, id) = vtb + vpb;
            ADJ<D>(a, L::A + k, id) = vtb * dt + vpb * 0.5 * dt;
        }
        const double upb = ADJ<D>(a, L::U, id);
        ADJ<D>(a, L::U, id) = ubF + upb;
        ADJ<D>(a, L::DU, id) = ubF * dt + upb * 0.5 * dt;
    } else {
#pragma unroll
        for(int k = 0; k < D; ++k) {
            ADJ<D>(a, L::X + k, id) += W(a, W_XB + k, i);
            ADJ<D>(a, L::V + k, id) += W(a, W_VB + k, i);
            ADJ<D>(a, L::A + k, id) = 0.0;
        }
        ADJ<D>(a, L::U, id) += ubF;
        ADJ<D>(a, L::DU, id) = 0.0;
    }
}

template <int D, int K, class T> static void launch_passes_t(Ctx &c, const AdjArgs &a, const BoxDev &box)
{
    const int g = grid_for(c.n);
    hipLaunchKernelGGL((adj_seed_kernel<D>), dim3(g), dim3(BLOCK), 0, c.stream, c.s, (int)c.n, a);
    hipLaunchKernelGGL((adj_force_kernel<D, K, T>), dim3(g), dim3(BLOCK), 0, c.stream, c.s, box, (int)c.n, a);
    hipLaunchKernelGGL((adj_pre_kernel<D, K, T>), dim3(g), dim3(BLOCK), 0, c.stream, c.s, box, (int)c.n, a);
    hipLaunchKernelGGL((adj_sum_kernel<D, K, T>), dim3(g), dim3(BLOCK), 0, c.stream, c.s, box, (int)c.n, a);
    hipLaunchKernelGGL((adj_state_kernel<D>), dim3(g), dim3(BLOCK), 0, c.stream, c.s, (int)c.n, a);
}

template <int D, int K> static void launch_passes(Ctx &c, const AdjArgs &a, const BoxDev &box)
{
    if(c.p.precision == SPHGPU_MIXED) launch_passes_t<D, K, float>(c, a, box);
    else launch_passes_t<D, K, double>(c, a, box);
}

int adjoint_step(Ctx &c, bool kick)
{
    AdjArgs a{};
    a.list = c.list;
    a.cnt = c.list_cnt;
    a.cnt2 = c.list_cnt2;
    a.cap = c.cap;
    a.w = c.ro.work;
    a.n = c.ro.n;
    a.adj = c.ro.adj;
    a.gamma = c.p.gamma;
    a.nngb = c.p.neighbor_number;
    a.balsara = c.p.av_use_balsara_switch;
    a.dt = c.ro.dt;
    a.kick = kick ? 1 : 0;
    const BoxDev box = make_box(c);
    const bool wend = c.p.kernel == SPHGPU_WENDLAND;
    switch(c.D) {
    case 1: launch_passes<1, 0>(c, a, box); break;
    case 2: wend ? launch_passes<2, 1>(c, a, box) : launch_passes<2, 0>(c, a, box); break;
    default: wend ? launch_passes<3, 1>(c, a, box) : launch_passes<3, 0>(c, a, box); break;
    }
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

void adjoint_free(Ctx &c)
{
    if(c.ro.ck) (void)hipFree(c.ro.ck);
    if(c.ro.ck_id) (void)hipFree(c.ro.ck_id);
    if(c.ro.adj) (void)hipFree(c.ro.adj);
    if(c.ro.work) (void)hipFree(c.ro.work);
    c.ro = Ctx::Rollout{};
}

// slots: 0 = state before Solver::initialize, 1..K = state at the start of
// step 0..K-1, K+1 = final state
int adjoint_alloc(Ctx &c, int steps)
{
    const int slots = steps + 2;
    if(c.ro.ck && c.ro.n == c.n && c.ro.slots >= slots) {
        c.ro.steps = 0;
        return SPHGPU_OK;
    }
    adjoint_free(c);
    int f[32];
    const int nck = ck_fields(c.D, f);
    const size_t n = (size_t)c.n;
    SPHGPU_HIP(hipMalloc(&c.ro.ck, (size_t)slots * nck * n * sizeof(double)));
    SPHGPU_HIP(hipMalloc(&c.ro.ck_id, (size_t)slots * n * sizeof(int)));
    SPHGPU_HIP(hipMalloc(&c.ro.adj, (size_t)(3 * c.D + 2) * n * sizeof(double)));
    SPHGPU_HIP(hipMalloc(&c.ro.work, (size_t)NWORK * n * sizeof(double)));
    c.ro.n = c.n;
    c.ro.slots = slots;
    c.ro.steps = 0;
    return SPHGPU_OK;
}

int adjoint_save(Ctx &c, int slot)
{
    int f[32];
    const int nck = ck_fields(c.D, f);
    const size_t n = (size_t)c.n;
    for(int k = 0; k < nck; ++k)
        SPHGPU_HIP(hipMemcpyAsync(c.ro.ck + ((size_t)slot * nck + k) * n, c.s.f[f[k]], n * sizeof(double),
                                  hipMemcpyDeviceToDevice, c.stream));
    SPHGPU_HIP(hipMemcpyAsync(c.ro.ck_id + (size_t)slot * n, c.s.id, n * sizeof(int), hipMemcpyDeviceToDevice, c.stream));
    return SPHGPU_OK;
}

int adjoint_restore(Ctx &c, int slot)
{
    int f[32];
    const int nck = ck_fields(c.D, f);
    const size_t n = (size_t)c.n;
    for(int k = 0; k < nck; ++k)
        SPHGPU_HIP(hipMemcpyAsync(c.s.f[f[k]], c.ro.ck + ((size_t)slot * nck + k) * n, n * sizeof(double),
                                  hipMemcpyDeviceToDevice, c.stream));
    SPHGPU_HIP(hipMemcpyAsync(c.s.id, c.ro.ck_id + (size_t)slot * n, n * sizeof(int), hipMemcpyDeviceToDevice, c.stream));
    return SPHGPU_OK;
}

// host seeds (caller order, n*D for vectors; NULL = 0) into the adjoint state
int adjoint_seed(Ctx &c, const double *dx, const double *dv, const double *du)
{
    const int D = c.D;
    const size_t n = (size_t)c.n;
    std::vector<double> h((size_t)(3 * D + 2) * n, 0.0);
    for(size_t i = 0; i < n; ++i)
        for(int k = 0; k < D; ++k) {
            if(dx) h[(size_t)k * n + i] = dx[i * D + k];
            if(dv) h[(size_t)(D + k) * n + i] = dv[i * D + k];
        }
    if(du)
        for(size_t i = 0; i < n; ++i) h[(size_t)2 * D * n + i] = du[i];
    SPHGPU_HIP(hipMemcpyAsync(c.ro.adj, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    return SPHGPU_OK;
}

int adjoint_fetch(Ctx &c, double *dx, double *dv, double *du)
{
    const int D = c.D;
    const size_t n = (size_t)c.n;
    std::vector<double> h((size_t)(2 * D + 1) * n);
    SPHGPU_HIP(hipMemcpyAsync(h.data(), c.ro.adj, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    for(size_t i = 0; i < n; ++i)
        for(int k = 0; k < D; ++k) {
            if(dx) dx[i * D + k] = h[(size_t)k * n + i];
            if(dv) dv[i * D + k] = h[(size_t)(D + k) * n + i];
        }
    if(du)
        for(size_t i = 0; i < n; ++i) du[i] = h[(size_t)2 * D * n + i];
    return SPHGPU_OK;
}

}  // namespace sphgpu
