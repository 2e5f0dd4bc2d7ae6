// Periodic boundary of the reference (include/periodic.hpp:12-72) on host
// vectors: minimum-image separation and the one-period wrap, with the
// device's comparison order.
#pragma once

#include <cmath>
#include <memory>

#include "sph/parameters.hpp"
#include "sph/vector_type.hpp"

namespace sph {

class Periodic {
public:
    Periodic() = default;
    void initialize(std::shared_ptr<SPHParameters> param) {
        m_is_valid = param->periodic.is_valid;
        m_dim = param->dim;
        if (!m_is_valid) return;
        for (int k = 0; k < m_dim; ++k) {
            m_max[k] = param->periodic.range_max[k];
            m_min[k] = param->periodic.range_min[k];
            m_range[k] = m_max[k] - m_min[k];
        }
    }
    // r_i - r_j, per axis the shortest of d, d + L, d - L
    vec_t calc_r_ij(const vec_t &r_i, const vec_t &r_j) const {
        vec_t d = r_i - r_j;
        if (!m_is_valid) return d;
        for (int k = 0; k < m_dim; ++k) {
            const real d1 = d[k], d2 = d1 + m_range[k], d3 = d1 - m_range[k];
            const real a1 = std::fabs(d1), a2 = std::fabs(d2), a3 = std::fabs(d3);
            real r = d3;
            if (a2 <= a3 && a2 <= a1) r = d2;
            if (a1 <= a2 && a1 <= a3) r = d1;
            d[k] = r;
        }
        return d;
    }
    void apply(vec_t &r) const {
        if (!m_is_valid) return;
        for (int k = 0; k < m_dim; ++k) {
            if (r[k] < m_min[k]) r[k] += m_range[k];
            else if (r[k] > m_max[k]) r[k] -= m_range[k];
        }
    }

private:
    bool m_is_valid = false;
    int m_dim = 3;
    vec_t m_max, m_min, m_range;
};

}  // namespace sph
