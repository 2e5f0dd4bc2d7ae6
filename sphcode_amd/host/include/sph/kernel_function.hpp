// Host evaluation of the SPH kernels behind the reference's KernelFunction
// interface (include/kernel/kernel_function.hpp:18-23): W(r, h), grad W and
// dW/dh for the cubic spline with support h (H = h/2, q = r/H < 2,
// include/kernel/cubic_spline.hpp:21-52) and Wendland C4 with support h
// (q = r/h < 1, include/kernel/wendland_kernel.hpp:23-49).  The device
// kernels evaluate the same formulas per pair; these serve host code that
// works on Simulation::get_particles().
#pragma once

#include <algorithm>
#include <memory>

#include "sph/exception.hpp"
#include "sph/parameters.hpp"
#include "sph/vector_type.hpp"

namespace sph {

constexpr real kPi = 3.14159265358979323846;

class KernelFunction {
public:
    virtual ~KernelFunction() = default;
    virtual real w(const real r, const real h) const = 0;                      // W(r,h)
    virtual vec_t dw(const vec_t &rij, const real r, const real h) const = 0;  // grad W(r,h)
    virtual real dhw(const real r, const real h) const = 0;                    // dW(r,h)/dh
};

inline real pow_dim(int dim, real h) { return dim == 1 ? h : dim == 2 ? h * h : h * h * h; }

class CubicSpline : public KernelFunction {
public:
    explicit CubicSpline(int dim)
        : m_dim(dim), m_sigma(dim == 1 ? 2.0 / 3.0 : dim == 2 ? 10.0 / (7.0 * kPi) : 1.0 / kPi) {}
    real w(const real r, const real h) const override {
        const real hinv = 2.0 / h, q = r * hinv;
        const real a = std::max(2.0 - q, 0.0), b = std::max(1.0 - q, 0.0);
        return m_sigma * pow_dim(m_dim, hinv) * (0.25 * a * a * a - b * b * b);
    }
    vec_t dw(const vec_t &rij, const real r, const real h) const override {
        if (r == 0.0) return vec_t();
        const real hinv = 2.0 / h, q = r * hinv;
        const real a = std::max(2.0 - q, 0.0), b = std::max(1.0 - q, 0.0);
        const real c = -m_sigma * pow_dim(m_dim, hinv) * hinv / r * (0.75 * a * a - 3.0 * b * b);
        return rij * c;
    }
    real dhw(const real r, const real h) const override {
        const real hinv = 2.0 / h, q = r * hinv;
        const real a = std::max(2.0 - q, 0.0), b = std::max(1.0 - q, 0.0);
        const real d = m_dim;
        return 0.5 * m_sigma * pow_dim(m_dim, hinv) * hinv *
               (a * a * ((3.0 + d) * 0.25 * q - 0.5 * d) + b * b * ((-3.0 - d) * q + d));
    }

private:
    int m_dim;
    real m_sigma;
};

class WendlandC4 : public KernelFunction {
public:
    explicit WendlandC4(int dim) : m_dim(dim), m_sigma(dim == 2 ? 9.0 / kPi : 495.0 / (32.0 * kPi)) {
        // the reference's C4 kernel is defined for DIM 2 and 3 only
        // (include/kernel/wendland_kernel.hpp:27)
        if (dim == 1) throw SPHException("Wendland C4 kernel is not available for DIM = 1");
    }
    real w(const real r, const real h) const override {
        const real q = r / h, t = std::max(1.0 - q, 0.0), t2 = t * t;
        return m_sigma / pow_dim(m_dim, h) * t2 * t2 * t2 * (1.0 + 6.0 * q + 35.0 / 3.0 * q * q);
    }
    vec_t dw(const vec_t &rij, const real r, const real h) const override {
        const real q = r / h, t = std::max(1.0 - q, 0.0), t2 = t * t;
        const real c = -56.0 / 3.0 * m_sigma / (pow_dim(m_dim, h) * h * h) * t2 * t2 * t * (1.0 + 5.0 * q);
        return rij * c;
    }
    real dhw(const real r, const real h) const override {
        const real q = r / h, t = std::max(1.0 - q, 0.0), t2 = t * t;
        const real d = m_dim;
        return -m_sigma / (3.0 * pow_dim(m_dim, h) * h) * t2 * t2 * t *
               (3.0 * d + 15.0 * d * q + (-56.0 + 17.0 * d) * q * q - 35.0 * (8.0 + d) * q * q * q);
    }

private:
    int m_dim;
    real m_sigma;
};

inline std::shared_ptr<KernelFunction> make_kernel(const SPHParameters &p) {
    if (p.kernel == KernelType::WENDLAND) return std::make_shared<WendlandC4>(p.dim);
    return std::make_shared<CubicSpline>(p.dim);
}

}  // namespace sph
