// Derivatives of the smoothing kernels with respect to r and h, for the
// adjoint of a step.  Both kernels have the form
//     W(r, h) = C h^-D phi(q),   q = kappa r / h,
// with kappa = 2, C = sigma_D 2^D for the cubic spline (support h, H = h/2;
// include/kernel/cubic_spline.hpp:21-52) and kappa = 1, C = sigma_D for
// Wendland C4 (include/kernel/wendland_kernel.hpp:23-49).  With
//     psi = phi'/q,  chi = -D phi - q phi',
// every derivative the adjoint needs is a polynomial in q times a power of h:
//     W     = C h^-D phi             dW/dh   = C h^-(D+1) chi
//     f     = (dW/dr)/r = C k^2 h^-(D+2) psi      (grad W = r_ij f)
//     df/dr = C k^3 h^-(D+3) psi'
//     df/dh = -C k^2 h^-(D+3) ((D+2) psi + q psi')
//     d2W/dh2  = -C h^-(D+2) ((D+1) chi + q chi'),  chi' = -(D+1) phi' - q phi''
//     d2W/dhdr = C k h^-(D+2) chi'
// Usable from host and device code.
#pragma once

#if defined(__HIPCC__)
#define KD_FN __host__ __device__ __forceinline__
#else
#define KD_FN inline
#endif

namespace sphgpu {

template <class T> struct KernelDerivs {
    T w, wh, whh, whr, f, fr, fh;
};

// phi, phi', phi'', psi, psi' of the cubic spline at q = 2r/h
template <class T> KD_FN void cubic_shape(T q, T &phi, T &d1, T &d2, T &psi, T &dpsi)
{
    if(q < T(1)) {
        phi = T(1) - T(1.5) * q * q + T(0.75) * q * q * q;
        d1 = T(-3) * q + T(2.25) * q * q;
        d2 = T(-3) + T(4.5) * q;
        psi = T(-3) + T(2.25) * q;
        dpsi = T(2.25);
    } else if(q < T(2)) {
        const T t = T(2) - q;
        phi = T(0.25) * t * t * t;
        d1 = T(-0.75) * t * t;
        d2 = T(1.5) * t;
        psi = d1 / q;
        dpsi = (d2 - psi) / q;
    } else {
        phi = d1 = d2 = psi = dpsi = T(0);
    }
}

// the same for Wendland C4 at q = r/h
template <class T> KD_FN void wendland_shape(T q, T &phi, T &d1, T &d2, T &psi, T &dpsi)
{
    if(q < T(1)) {
        const T t = T(1) - q;
        const T t2 = t * t, t4 = t2 * t2;
        phi = t4 * t2 * (T(1) + T(6) * q + T(35.0 / 3.0) * q * q);
        psi = T(-56.0 / 3.0) * t4 * t * (T(1) + T(5) * q);
        dpsi = T(560) * q * t4;
        d1 = q * psi;
        d2 = psi + q * dpsi;
    } else {
        phi = d1 = d2 = psi = dpsi = T(0);
    }
}

template <int D> struct KernelConst {
    static constexpr double pi = 3.14159265358979323846;
    static constexpr double cubic = D == 1 ? 2.0 / 3.0 * 2.0 : D == 2 ? 10.0 / (7.0 * pi) * 4.0 : 1.0 / pi * 8.0;
    static constexpr double wendland = D == 1 ? 0.0 : D == 2 ? 9.0 / pi : 495.0 / (32.0 * pi);
};

template <class T, int D> KD_FN T ipow(T x)
{
    T y = T(1);
    for(int k = 0; k < D; ++k) y *= x;
    return y;
}

// K = 0: cubic spline, K = 1: Wendland C4
template <int D, int K, class T> KD_FN KernelDerivs<T> kernel_derivs(T r, T h)
{
    constexpr T kappa = K == 0 ? T(2) : T(1);
    const T C = K == 0 ? T(KernelConst<D>::cubic) : T(KernelConst<D>::wendland);
    const T hinv = T(1) / h;
    const T q = kappa * r * hinv;
    T phi, d1, d2, psi, dpsi;
    if constexpr(K == 0) cubic_shape<T>(q, phi, d1, d2, psi, dpsi);
    else wendland_shape<T>(q, phi, d1, d2, psi, dpsi);
    const T hD = C * ipow<T, D>(hinv);  // C h^-D
    const T h1 = hD * hinv, h2 = h1 * hinv, h3 = h2 * hinv;
    const T chi = -T(D) * phi - q * d1;
    const T dchi = -T(D + 1) * d1 - q * d2;
    KernelDerivs<T> o;
    o.w = hD * phi;
    o.wh = h1 * chi;
    o.whh = -h2 * (T(D + 1) * chi + q * dchi);
    o.whr = kappa * h2 * dchi;
    o.f = kappa * kappa * h2 * psi;
    o.fr = kappa * kappa * kappa * h3 * dpsi;
    o.fh = -kappa * kappa * h3 * (T(D + 2) * psi + q * dpsi);
    return o;
}

}  // namespace sphgpu
