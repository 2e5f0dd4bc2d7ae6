// Device-side primitives shared by every stage: wave64 helpers, periodic
// minimum image, SPH kernels and the Hernquist-Katz gravitational softening.
//
// Distances that decide set membership (neighbour candidates, density
// cut-off, force cut-off, cell pruning) are evaluated in the reference's
// operation order with FP contraction disabled, so FP64 sets match the
// reference's exactly.  Sums over pairs may be contracted.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sphgpu {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;

// ------------------------------------------------------------- wave64 helpers
__device__ __forceinline__ int wave_lane() { return __lane_id(); }

// Mask of the lanes with p set, in scalar registers (the compare result
// itself, with no round trip of p through a vector register).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ bool wave_any(bool p) { return ballot(p) != 0ull; }

// Loads and stores of data a wave streams (candidate tiles), marked
// non-temporal so that they do not push reused neighbour records out of L2
// (measured: 11 % fewer bytes fetched by the pre-interaction's pair pass,
// 9 % by the force pass).
template <class T> __device__ __forceinline__ T stream_ld(const T *p) { return __builtin_nontemporal_load(p); }
template <class T> __device__ __forceinline__ void stream_st(T *p, T v) { __builtin_nontemporal_store(v, p); }

// This lane's bit of a wave-uniform lane mask, as a condition the vector
// instructions take straight from the scalar registers.
__device__ __forceinline__ bool lane_in(uint64_t mask) { return __builtin_amdgcn_inverse_ballot_w64(mask); }

// 64 x 64 bit-matrix transpose across the wave: lane l returns the word
// whose bit k is bit l of lane k's word.  Six butterfly stages: at block
// size s, lanes l and l ^ s swap the s-bit blocks that lie across the
// diagonal.  Every lane of the wave must take part.
__device__ __forceinline__ uint64_t wave_transpose64(uint64_t a)
{
    const int lane = wave_lane();
    constexpr uint64_t M[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                               0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for(int st = 0; st < 6; ++st) {
        const int s = 32 >> st;
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)a, s, WAVE);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(a >> 32), s, WAVE);
        const uint64_t p = ((uint64_t)hi << 32) | lo;
        const bool up = (lane & s) != 0;
        const uint64_t t = (((up ? p : a) >> s) ^ (up ? a : p)) & M[st];
        a ^= up ? t : (t << s);
    }
    return a;
}

__device__ __forceinline__ int wave_min_int(int v)
{
    for(int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off, WAVE);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uniform64(uint64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// A wave-uniform double moved to scalar registers.
__device__ __forceinline__ double uniform_d(double v)
{
    return __longlong_as_double((long long)uniform64((uint64_t)__double_as_longlong(v)));
}

// MI355X deals workgroups round-robin over its 8 XCDs, each with its own L2.
// Renumbering the launch so that XCD x runs one contiguous range of
// logical blocks keeps spatially adjacent particles (adjacent in Morton
// order) and their neighbours in the same L2.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    const int per = nb / NUM_XCD, rem = nb % NUM_XCD;
    const int x = b % NUM_XCD, k = b / NUM_XCD;
    return x * per + (x < rem ? x : rem) + k;
}
__device__ __forceinline__ int xcd_block() { return uniform(xcd_block((int)blockIdx.x, (int)gridDim.x)); }

// Persistent waves: tasks [0, ntask) are split into NUM_XCD contiguous
// shares, and the waves of each XCD take the tasks of its share one at a
// time from that XCD's counter (hardware deals block b to XCD b mod 8).
// A wave whose share is used up moves on to the next XCD's share, so no
// XCD idles while another still has tasks; after visiting all NUM_XCD
// shares it gets -1.  Call with all lanes active.
struct TaskShare {
    int *work;
    int ntask;
    int x, visited = 0;
    int lo, hi;
    __device__ __forceinline__ void set(int xcd)
    {
        x = xcd;
        lo = (int)((int64_t)ntask * x / NUM_XCD);
        hi = (int)((int64_t)ntask * (x + 1) / NUM_XCD);
    }
    __device__ __forceinline__ TaskShare(int *work_, int ntask_) : work(work_), ntask(ntask_)
    {
        set((int)(blockIdx.x % NUM_XCD));
    }
    __device__ __forceinline__ int next()
    {
        while(true) {
            int t = 0;
            if(wave_lane() == 0) t = atomicAdd(work + x, 1);
            t = lo + __builtin_amdgcn_readfirstlane(t);
            if(t < hi) return t;
            if(++visited == NUM_XCD) return -1;
            set(x + 1 < NUM_XCD ? x + 1 : 0);
        }
    }
};

// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <class T> __device__ __forceinline__ T wave_sum(T v)
{
    for(int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

template <class T> __device__ __forceinline__ T wave_minv(T v)
{
    for(int off = 32; off > 0; off >>= 1) {
        const T o = __shfl_xor(v, off, WAVE);
        v = o < v ? o : v;
    }
    return v;
}

template <class T> __device__ __forceinline__ T wave_maxv(T v)
{
    for(int off = 32; off > 0; off >>= 1) {
        const T o = __shfl_xor(v, off, WAVE);
        v = o > v ? o : v;
    }
    return v;
}

// Inclusive prefix maximum over the 64 lanes of non-negative values, by
// DPP: row shifts by 1, 2, 4, 8 within each row of 16 lanes, then lane 15's
// value into row 1 and lane 47's into row 3, then lane 31's into rows 2 and
// 3.  Lanes whose DPP source is outside the pattern read 0, the identity.
__device__ __forceinline__ int wave_incl_max_nonneg(int x)
{
    int y;
    y = __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x = y > x ? y : x;
    y = __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x = y > x ? y : x;
    y = __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x = y > x ? y : x;
    y = __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x = y > x ? y : x;
    y = __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x = y > x ? y : x;
    y = __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    x = y > x ? y : x;
    return x;
}

// Lane l receives lane l-1's value; lane 0 receives 0 (DPP wave_shr:1).
__device__ __forceinline__ int wave_shr1(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, false); }

// Read of a record the kernel never writes, through the constant address
// space: a wave-uniform address then becomes a scalar load even where the
// kernel also stores through other pointers (which otherwise forces vector
// loads).  The host compilation pass only needs a plain read.
template <class T> __device__ __forceinline__ T load_ro(const T *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    // word by word: a whole-struct copy is split by field type, and some
    // pieces were still selected as vector loads
    static_assert(sizeof(T) % 16 == 0 && alignof(T) >= 16, "load_ro reads 16-byte words");
    T out;
    const __attribute__((address_space(4))) uint4 *q = (const __attribute__((address_space(4))) uint4 *)p;
    uint4 *o = reinterpret_cast<uint4 *>(&out);
#pragma unroll
    for(int k = 0; k < (int)(sizeof(T) / 16); ++k) o[k] = q[k];
    return out;
#else
    return *p;
#endif
}

// |r| in FP32 of a gather candidate from its FP64 r^2 (mixed mode): the
// hardware FP64 square root (one instruction, no refinement sequence)
// rounded to FP32, which is the correctly rounded FP32 |r| except when the
// FP64 result falls within its error of an FP32 rounding boundary.  The
// search stores it and the adjoint recomputes it with this same function.
__device__ __forceinline__ float gather_r_f32(double r2) { return (float)__builtin_amdgcn_sqrt(r2); }

// --------------------------------------------------------------- scalar math
template <class T> __host__ __device__ __forceinline__ T sq(T x) { return x * x; }
template <class T> __host__ __device__ __forceinline__ T cube(T x) { return x * x * x; }

// h^D (include/kernel/kernel_function.hpp:8-16)
template <int D, class T> __device__ __forceinline__ T pow_dim(T h)
{
    if constexpr(D == 1) return h;
    else if constexpr(D == 2) return h * h;
    else return h * h * h;
}

// h^(D-1) (src/pre_interaction.cpp:217-225)
template <int D, class T> __device__ __forceinline__ T pow_dim1(T h)
{
    if constexpr(D == 1) return T(1);
    else if constexpr(D == 2) return h;
    else return h * h;
}

// x^(1/D); the reference's -ffast-math build evaluates std::pow(x, 1.0/DIM)
// (src/pre_interaction.cpp:64) as x, sqrt(x), cbrt(x).
template <int D> __device__ __forceinline__ double root_dim(double x)
{
    if constexpr(D == 1) return x;
    else if constexpr(D == 2) return sqrt(x);
    else return cbrt(x);
}

// Unit-sphere volume A_D (src/pre_interaction.cpp:61-63).
template <int D> struct UnitVolume {
    static constexpr double value = D == 1 ? 2.0 : D == 2 ? 3.14159265358979323846 : 4.0 * 3.14159265358979323846 / 3.0;
};

// The gather search radius sml = h_guess * ratio, h_guess = (N_ngb m /
// (rho A_D))^(1/D) (src/pre_interaction.cpp:61-64).
template <int D> __device__ __forceinline__ double sml_guess(double m, double rho, double nngb, double ratio)
{
    return root_dim<D>(nngb * m / (rho * UnitVolume<D>::value)) * ratio;
}

// -------------------------------------------------------- periodic boundary
// Periodic::calc_r_ij (include/periodic.hpp:34-59): per axis, the smallest
// of |d|, |d+L|, |d-L| in the reference's tie order.
struct BoxDev {
    int on;
    double lo[3], hi[3], len[3];
};

template <int D> __device__ __forceinline__ void min_image(const BoxDev &b, double d[D])
{
#pragma clang fp contract(off)
    if(!b.on) return;
#pragma unroll
    for(int k = 0; k < D; ++k) {
        const double d1 = d[k];
        const double d2 = d1 + b.len[k];
        const double d3 = d1 - b.len[k];
        const double a1 = fabs(d1), a2 = fabs(d2), a3 = fabs(d3);
        double r = d3;
        if(a2 <= a3 && a2 <= a1) r = d2;
        if(a1 <= a2 && a1 <= a3) r = d1;
        d[k] = r;
    }
}

// r_ij = x_i - x_j with minimum image, and its squared norm summed left to
// right without contraction (include/vector_type.hpp:194-227).
template <int D> __device__ __forceinline__ double sep2(const BoxDev &b, const double xi[D], const double xj[D], double d[D])
{
#pragma clang fp contract(off)
#pragma unroll
    for(int k = 0; k < D; ++k) d[k] = xi[k] - xj[k];
    min_image<D>(b, d);
    double r2 = d[0] * d[0];
#pragma unroll
    for(int k = 1; k < D; ++k) r2 = r2 + d[k] * d[k];
    return r2;
}

// Periodic::apply (include/periodic.hpp:61-72): one-period wrap.
template <int D> __device__ __forceinline__ void wrap(const BoxDev &b, double x[D])
{
#pragma clang fp contract(off)
    if(!b.on) return;
#pragma unroll
    for(int k = 0; k < D; ++k) {
        if(x[k] < b.lo[k]) x[k] += b.len[k];
        else if(x[k] > b.hi[k]) x[k] -= b.len[k];
    }
}

// ------------------------------------------------------------- SPH kernels
// Both kernels are evaluated per pair from q and per-particle constants of
// h: for W, grad W = r_ij * coef and dW/dh.  Formulas follow
// include/kernel/cubic_spline.hpp:21-52 and
// include/kernel/wendland_kernel.hpp:23-49; the per-h constants are
// hoisted out of the pair loop.

// Cubic spline with support h: H = h/2, q = r/H < 2.
// 1/sqrt(x) in the pair arithmetic type (HIP's rsqrt takes double only, so
// a float argument would be widened and refined in FP64).  The FP32 form is
// the hardware reciprocal square root (about 1 ulp) without the rescaling
// of denormal inputs: the separations it is applied to are far above
// FLT_MIN (the caller excludes zero).
template <class T> __device__ __forceinline__ T rsqrt_t(T x)
{
    if constexpr(sizeof(T) == 4) return __builtin_amdgcn_rsqf(x);
    else return rsqrt(x);
}

template <int D, class T> struct Cubic {
    static constexpr double sigma_d = D == 1 ? 2.0 / 3.0 : D == 2 ? 10.0 / (7.0 * 3.14159265358979323846) : 1.0 / 3.14159265358979323846;
    T qscale;   // 1/H
    T wnorm;    // sigma / H^D
    T gnorm;    // sigma / (H^D H)
    __device__ __forceinline__ explicit Cubic(T h) : Cubic(h, T(1) / h) {}
    // hinv = 1/h supplied by the caller
    __device__ __forceinline__ Cubic(T /*h*/, T hinv)
    {
        qscale = T(2) * hinv;
        wnorm = T(sigma_d) * pow_dim<D>(qscale);
        gnorm = wnorm * qscale;
    }
    __device__ __forceinline__ T q_of(T r) const { return r * qscale; }
    __device__ __forceinline__ T w(T q) const
    {
        const T a = fmax(T(2) - q, T(0)), b = fmax(T(1) - q, T(0));
        return wnorm * (T(0.25) * a * a * a - b * b * b);
    }
    // grad W = r_ij * dw_coef(q, r)
    __device__ __forceinline__ T dw_coef(T q, T r) const
    {
        if(r == T(0)) return T(0);
        const T a = fmax(T(2) - q, T(0)), b = fmax(T(1) - q, T(0));
        return -gnorm / r * (T(0.75) * a * a - T(3) * b * b);
    }
    __device__ __forceinline__ T dhw(T q) const
    {
        const T a = fmax(T(2) - q, T(0)), b = fmax(T(1) - q, T(0));
        return T(0.5) * gnorm * (a * a * (T((3.0 + D) * 0.25) * q - T(0.5 * D)) + b * b * (T(-3.0 - D) * q + T(D)));
    }
};

// Wendland C4 with support h, q = r/h < 1 (Dehnen & Aly 2012).
template <int D, class T> struct WendlandC4 {
    static constexpr double sigma_d = D == 1 ? 0.0 : D == 2 ? 9.0 / 3.14159265358979323846 : 495.0 / (32.0 * 3.14159265358979323846);
    T qscale;   // 1/h
    T wnorm;    // sigma / h^D
    T gnorm;    // 56/3 sigma / (h^D h^2)
    T hnorm;    // sigma / (3 h^D h)
    __device__ __forceinline__ explicit WendlandC4(T h) : WendlandC4(h, T(1) / h) {}
    // hinv = 1/h supplied by the caller
    __device__ __forceinline__ WendlandC4(T /*h*/, T hinv)
    {
        qscale = hinv;
        wnorm = T(sigma_d) * pow_dim<D>(hinv);
        gnorm = T(56.0 / 3.0) * wnorm * qscale * qscale;
        hnorm = wnorm * qscale * T(1.0 / 3.0);
    }
    __device__ __forceinline__ T q_of(T r) const { return r * qscale; }
    __device__ __forceinline__ T w(T q) const
    {
        const T t = fmax(T(1) - q, T(0));
        const T t2 = t * t;
        return wnorm * (t2 * t2 * t2) * (T(1) + T(6) * q + T(35.0 / 3.0) * q * q);
    }
    __device__ __forceinline__ T dw_coef(T q, T /*r*/) const
    {
        const T t = fmax(T(1) - q, T(0));
        const T t2 = t * t;
        return -gnorm * (t2 * t2 * t) * (T(1) + T(5) * q);
    }
    __device__ __forceinline__ T dhw(T q) const
    {
        const T t = fmax(T(1) - q, T(0));
        const T t2 = t * t;
        return -hnorm * (t2 * t2 * t) *
               (T(3 * D) + T(15 * D) * q + T(-56 + 17 * D) * q * q - T(35 * (8 + D)) * q * q * q);
    }
};

template <int D, int K, class T> struct KernelOf;
template <int D, class T> struct KernelOf<D, 0, T> { using type = Cubic<D, T>; };
template <int D, class T> struct KernelOf<D, 1, T> { using type = WendlandC4<D, T>; };

// ------------------------------------------- Hernquist & Katz softening
// f = potential kernel and g = force kernel of src/bhtree.cpp:273-299 with
// e = h/2, u = r/e; g(r) * r_ij is the softened r_ij / r^3.
// Both softened kernels at once, without branches, given u = r/e, 1/e and
// 1/r (1/r may be 0 at r = 0, where only the inner piece is used); the
// polynomials of src/bhtree.cpp:273-299 in Horner form:
//   inner  f = (1.4 + u^2 (-1/6 + u^2 (3/40 - u/40))) / e
//          g = (4/3 + u^2 (-6/5 + u/2)) / e^3
//   middle f = -1/(15 r) + (1.6 + u^2 (-4/3 + u (1 + u (-3/10 + u/30)))) / e
//          g = (-1/15 + u^3 (8/3 + u (-3 + u (6/5 - u/6)))) / r^3
template <class T> __device__ __forceinline__ void soft_both(T u, T einv, T rinv, T &pot, T &acc)
{
    const T u2 = u * u;
    const T u3 = u2 * u;
    const T e3 = einv * einv * einv;
    const T r3 = rinv * rinv * rinv;
    const T pot_in = (T(1.4) + u2 * (T(-1.0 / 6.0) + u2 * (T(3.0 / 40.0) - u * T(1.0 / 40.0)))) * einv;
    const T acc_in = (T(4.0 / 3.0) + u2 * (T(-1.2) + u * T(0.5))) * e3;
    const T pot_mid = T(-1.0 / 15.0) * rinv + (T(1.6) + u2 * (T(-4.0 / 3.0) + u * (T(1) + u * (T(-0.3) + u * T(1.0 / 30.0))))) * einv;
    const T acc_mid = (T(-1.0 / 15.0) + u3 * (T(8.0 / 3.0) + u * (T(-3) + u * (T(1.2) - u * T(1.0 / 6.0))))) * r3;
    const bool in = u < T(1), mid = u < T(2);
    pot = in ? pot_in : (mid ? pot_mid : rinv);
    acc = in ? acc_in : (mid ? acc_mid : r3);
}


}  // namespace sphgpu
