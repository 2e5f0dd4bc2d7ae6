// vec_t of the reference (include/vector_type.hpp): a DIM-component vector
// with component access and vector / scalar arithmetic.  DIM is a run-time
// setting here (SPHParameters::dim), so vec_t holds three components and
// the ones past dim stay zero (every operation keeps them zero when its
// operands have them zero).
#pragma once

#include <cmath>

namespace sph {

using real = double;

struct vec_t {
    real vec[3] = {0.0, 0.0, 0.0};

    vec_t() = default;
    vec_t(real x, real y = 0.0, real z = 0.0) : vec{x, y, z} {}

    real &operator[](int i) { return vec[i]; }
    const real &operator[](int i) const { return vec[i]; }

    // assignment of a scalar sets every component (center = 0.0 in the
    // reference's tree)
    vec_t &operator=(real a) {
        vec[0] = vec[1] = vec[2] = a;
        return *this;
    }

    vec_t operator-() const { return vec_t(-vec[0], -vec[1], -vec[2]); }
    vec_t &operator+=(const vec_t &a) {
        for (int k = 0; k < 3; ++k) vec[k] += a.vec[k];
        return *this;
    }
    vec_t &operator-=(const vec_t &a) {
        for (int k = 0; k < 3; ++k) vec[k] -= a.vec[k];
        return *this;
    }
    vec_t &operator*=(real a) {
        for (real &x : vec) x *= a;
        return *this;
    }
    vec_t &operator/=(real a) {
        for (real &x : vec) x /= a;
        return *this;
    }
    vec_t operator+(const vec_t &a) const { return vec_t(vec[0] + a[0], vec[1] + a[1], vec[2] + a[2]); }
    vec_t operator-(const vec_t &a) const { return vec_t(vec[0] - a[0], vec[1] - a[1], vec[2] - a[2]); }
    vec_t operator*(real a) const { return vec_t(vec[0] * a, vec[1] * a, vec[2] * a); }
    vec_t operator/(real a) const { return vec_t(vec[0] / a, vec[1] / a, vec[2] / a); }
};

inline vec_t operator*(real a, const vec_t &v) { return v * a; }

inline real inner_product(const vec_t &a, const vec_t &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline real abs2(const vec_t &a) { return inner_product(a, a); }
inline real abs(const vec_t &a) { return std::sqrt(abs2(a)); }
inline real distance(const vec_t &a, const vec_t &b) { return abs(a - b); }

}  // namespace sph
