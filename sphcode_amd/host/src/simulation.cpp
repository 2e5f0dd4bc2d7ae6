// Simulation over a sphgpu context (include/simulation.hpp:24-45,
// src/simulation.cpp:12-76).

#include <cstring>

#include "sph/exception.hpp"
#include "sph/simulation.hpp"

namespace sph {

void check_status(int status, const char *call) {
    if (status != SPHGPU_OK) throw SPHException(std::string(call) + ": " + sphgpu_last_error(), status);
}

Simulation::Simulation(std::shared_ptr<SPHParameters> param) : m_param(std::move(param)) {
    const sphgpu_params c = to_c_params(*m_param);
    // the context starts at time.start with dt = 0 (src/simulation.cpp:26-29)
    check_status(sphgpu_create(&c, &m_ctx), "sphgpu_create");
}

Simulation::~Simulation() {
    if (m_ctx) sphgpu_destroy(m_ctx);
}

namespace {
// SoA staging buffers for one transfer
struct Staging {
    std::vector<double> pos, vel, vel_p, acc, mass, dens, pres, ene, ene_p, dene, sml, sound, balsara, alpha, gradh,
        phi;
    std::vector<int32_t> id, neighbor;
    sphgpu_particles view{};

    Staging(size_t n, int dim)
        : pos(n * dim), vel(n * dim), vel_p(n * dim), acc(n * dim), mass(n), dens(n), pres(n), ene(n), ene_p(n),
          dene(n), sml(n), sound(n), balsara(n), alpha(n), gradh(n), phi(n), id(n), neighbor(n) {
        view.n = (int64_t)n;
        view.pos = pos.data();
        view.vel = vel.data();
        view.vel_p = vel_p.data();
        view.acc = acc.data();
        view.mass = mass.data();
        view.dens = dens.data();
        view.pres = pres.data();
        view.ene = ene.data();
        view.ene_p = ene_p.data();
        view.dene = dene.data();
        view.sml = sml.data();
        view.sound = sound.data();
        view.balsara = balsara.data();
        view.alpha = alpha.data();
        view.gradh = gradh.data();
        view.phi = phi.data();
        view.id = id.data();
        view.neighbor = neighbor.data();
    }
};
}  // namespace

void Simulation::set_particles(const std::vector<SPHParticle> &particles) {
    const int dim = m_param->dim;
    const size_t n = particles.size();
    if (n == 0) throw SPHException("set_particles: no particles");
    Staging s(n, dim);
    for (size_t i = 0; i < n; ++i) {
        const SPHParticle &p = particles[i];
        for (int k = 0; k < dim; ++k) {
            s.pos[i * dim + k] = p.pos[k];
            s.vel[i * dim + k] = p.vel[k];
            s.vel_p[i * dim + k] = p.vel_p[k];
            s.acc[i * dim + k] = p.acc[k];
        }
        s.mass[i] = p.mass;
        s.dens[i] = p.dens;
        s.pres[i] = p.pres;
        s.ene[i] = p.ene;
        s.ene_p[i] = p.ene_p;
        s.dene[i] = p.dene;
        s.sml[i] = p.sml;
        s.sound[i] = p.sound;
        s.balsara[i] = p.balsara;
        s.alpha[i] = p.alpha;
        s.gradh[i] = p.gradh;
        s.phi[i] = p.phi;
        s.id[i] = p.id;
        s.neighbor[i] = p.neighbor;
    }
    check_status(sphgpu_upload(m_ctx, &s.view), "sphgpu_upload");
    m_particles = particles;
}

void Simulation::sync_from_device() {
    const int dim = m_param->dim;
    const size_t n = m_particles.size();
    if (n == 0) return;
    Staging s(n, dim);
    check_status(sphgpu_download(m_ctx, &s.view), "sphgpu_download");
    for (size_t i = 0; i < n; ++i) {
        SPHParticle &p = m_particles[i];
        for (int k = 0; k < dim; ++k) {
            p.pos[k] = s.pos[i * dim + k];
            p.vel[k] = s.vel[i * dim + k];
            p.vel_p[k] = s.vel_p[i * dim + k];
            p.acc[k] = s.acc[i * dim + k];
        }
        p.mass = s.mass[i];
        p.dens = s.dens[i];
        p.pres = s.pres[i];
        p.ene = s.ene[i];
        p.ene_p = s.ene_p[i];
        p.dene = s.dene[i];
        p.sml = s.sml[i];
        p.sound = s.sound[i];
        p.balsara = s.balsara[i];
        p.alpha = s.alpha[i];
        p.gradh = s.gradh[i];
        p.phi = s.phi[i];
        p.id = s.id[i];
        p.neighbor = s.neighbor[i];
    }
}

double Simulation::get_time() const {
    double t, dt, hv;
    check_status(sphgpu_get_time(m_ctx, &t, &dt, &hv), "sphgpu_get_time");
    return t;
}

double Simulation::get_dt() const {
    double t, dt, hv;
    check_status(sphgpu_get_time(m_ctx, &t, &dt, &hv), "sphgpu_get_time");
    return dt;
}

double Simulation::get_h_per_v_sig() const {
    double t, dt, hv;
    check_status(sphgpu_get_time(m_ctx, &t, &dt, &hv), "sphgpu_get_time");
    return hv;
}

void Simulation::set_dt(double dt) { check_status(sphgpu_set_time(m_ctx, get_time(), dt), "sphgpu_set_time"); }

void Simulation::update_time() {
    double t, dt, hv;
    check_status(sphgpu_get_time(m_ctx, &t, &dt, &hv), "sphgpu_get_time");
    check_status(sphgpu_set_time(m_ctx, t + dt, dt), "sphgpu_set_time");
}

void Simulation::restart(double time, double dt, double h_per_v_sig) {
    check_status(sphgpu_restart(m_ctx, time, dt, h_per_v_sig), "sphgpu_restart");
}

void Simulation::make_tree() { check_status(sphgpu_make_tree(m_ctx), "sphgpu_make_tree"); }

std::array<double, 7> Simulation::energy() const {
    std::array<double, 7> e{};
    check_status(sphgpu_energy(m_ctx, e.data()), "sphgpu_energy");
    return e;
}

}  // namespace sph
