// Simulation over a sphgpu context (include/simulation.hpp:24-45,
// src/simulation.cpp:12-76), with the host copy of the particles kept in
// step with the device on demand.

#include <cstring>

#include "sph/exception.hpp"
#include "sph/simulation.hpp"

namespace sph {

void check_status(int status, const char *call) {
    if (status != SPHGPU_OK) throw SPHException(std::string(call) + ": " + sphgpu_last_error(), status);
}

void BHTree::resize(int num) {
    if (num != m_sim.get_particle_num())
        throw SPHException("BHTree::resize: the device tree is sized for " + std::to_string(m_sim.get_particle_num()) +
                           " particles, not " + std::to_string(num));
}

void BHTree::make(std::vector<SPHParticle> &particles, int num) {
    if (&particles != &m_sim.get_particles() || num != m_sim.get_particle_num())
        throw SPHException("BHTree::make: the tree is built from the simulation's own particles");
    m_sim.make_tree();
}

Simulation::Simulation(std::shared_ptr<SPHParameters> param) : m_param(std::move(param)) {
    const sphgpu_params c = to_c_params(*m_param);
    // the context starts at time.start with dt = 0 (src/simulation.cpp:26-29)
    check_status(sphgpu_create(&c, &m_ctx), "sphgpu_create");
    m_kernel = make_kernel(*m_param);
    m_periodic = std::make_shared<Periodic>();
    m_periodic->initialize(m_param);
    m_tree = std::make_shared<BHTree>(*this);
}

Simulation::~Simulation() {
    if (m_ctx) sphgpu_destroy(m_ctx);
}

namespace {
// SoA staging buffers for one transfer
struct Staging {
    std::vector<double> pos, vel, vel_p, acc, mass, dens, pres, ene, ene_p, dene, sml, sound, balsara, alpha, gradh,
        phi, gd, gp, gv;
    std::vector<int32_t> id, neighbor;
    sphgpu_particles view{};

    Staging(size_t n, int dim, bool grads)
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
        if (grads) {
            gd.resize(n * dim);
            gp.resize(n * dim);
            gv.resize(n * dim * dim);
            view.grad_density = gd.data();
            view.grad_pressure = gp.data();
            view.grad_velocity = gv.data();
        }
    }
};

const char *kGradVelocity[3] = {"grad_velocity_0", "grad_velocity_1", "grad_velocity_2"};
}  // namespace

void Simulation::set_particles(const std::vector<SPHParticle> &particles) {
    if (particles.empty()) throw SPHException("set_particles: no particles");
    m_particles = particles;
    m_particle_num = (int)particles.size();
    m_host_current = true;
    m_host_dirty = true;
    m_started = false;
    push();
}

void Simulation::set_particle_num(int num) {
    if (num != (int)m_particles.size() && !m_particles.empty())
        throw SPHException("set_particle_num: " + std::to_string(num) + " differs from the " +
                           std::to_string(m_particles.size()) + " particles set");
    m_particle_num = num;
}

// Uploads the host copy.  Mid-run (after a pre-interaction) the upload
// continues from the context's time, dt and min h/v_sig, as a restart does;
// before it, the next pre-interaction performs the initial smoothing.
void Simulation::push() {
    if (!m_host_dirty) return;
    const int dim = m_param->dim;
    const size_t n = m_particles.size();
    if (n == 0) throw SPHException("no particles set");
    double t = 0.0, dt = 0.0, hv = 0.0;
    if (m_started) check_status(sphgpu_get_time(m_ctx, &t, &dt, &hv), "sphgpu_get_time");
    const bool grads = m_vector_arrays.count("grad_density") != 0;
    Staging s(n, dim, grads);
    for (size_t i = 0; i < n; ++i) {
        const SPHParticle &p = m_particles[i];
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
    if (grads) {
        const auto &gd = m_vector_arrays["grad_density"];
        const auto &gp = m_vector_arrays["grad_pressure"];
        for (size_t i = 0; i < n; ++i)
            for (int l = 0; l < dim; ++l) {
                s.gd[i * dim + l] = gd[i][l];
                s.gp[i * dim + l] = gp[i][l];
                for (int k = 0; k < dim; ++k) s.gv[(i * dim + k) * dim + l] = m_vector_arrays[kGradVelocity[k]][i][l];
            }
    }
    check_status(sphgpu_upload(m_ctx, &s.view), "sphgpu_upload");
    if (m_started) {
        if (!(hv > 0.0)) throw SPHException("no min h/v_sig to continue from");
        check_status(sphgpu_restart(m_ctx, t, dt, hv), "sphgpu_restart");
    }
    m_host_dirty = false;
    m_host_current = true;
}

void Simulation::pull() {
    const int dim = m_param->dim;
    const size_t n = m_particles.size();
    if (n == 0) return;
    const bool grads = m_vector_arrays.count("grad_density") != 0;
    Staging s(n, dim, grads);
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
    if (grads) {
        auto &gd = m_vector_arrays["grad_density"];
        auto &gp = m_vector_arrays["grad_pressure"];
        for (size_t i = 0; i < n; ++i)
            for (int l = 0; l < dim; ++l) {
                gd[i][l] = s.gd[i * dim + l];
                gp[i][l] = s.gp[i * dim + l];
                for (int k = 0; k < dim; ++k) m_vector_arrays[kGradVelocity[k]][i][l] = s.gv[(i * dim + k) * dim + l];
            }
    }
    m_host_current = true;
}

std::vector<SPHParticle> &Simulation::get_particles() {
    if (!m_host_current) pull();
    m_host_dirty = true;
    return m_particles;
}

const std::vector<SPHParticle> &Simulation::particles_view() {
    if (!m_host_current) pull();
    return m_particles;
}

void Simulation::sync_from_device() {
    if (m_host_dirty) return;  // the host copy is the newer state
    pull();
}

void *Simulation::device() {
    push();
    m_host_current = false;
    return m_ctx;
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

void Simulation::set_time(double time) { check_status(sphgpu_set_time(m_ctx, time, get_dt()), "sphgpu_set_time"); }

void Simulation::set_dt(double dt) { check_status(sphgpu_set_time(m_ctx, get_time(), dt), "sphgpu_set_time"); }

void Simulation::set_h_per_v_sig(double h_per_v_sig) {
    push();
    check_status(sphgpu_restart(m_ctx, get_time(), get_dt(), h_per_v_sig), "sphgpu_restart");
    m_started = true;
    m_host_current = false;
}

void Simulation::update_time() {
    double t, dt, hv;
    check_status(sphgpu_get_time(m_ctx, &t, &dt, &hv), "sphgpu_get_time");
    check_status(sphgpu_set_time(m_ctx, t + dt, dt), "sphgpu_set_time");
}

void Simulation::restart(double time, double dt, double h_per_v_sig) {
    push();
    check_status(sphgpu_restart(m_ctx, time, dt, h_per_v_sig), "sphgpu_restart");
    m_started = true;
    m_host_current = false;
}

void Simulation::make_tree() { check_status(sphgpu_make_tree(device()), "sphgpu_make_tree"); }

void Simulation::add_scalar_array(const std::vector<std::string> &names) {
    for (const auto &name : names) m_scalar_arrays[name].assign(m_particles.size(), 0.0);
}

void Simulation::add_vector_array(const std::vector<std::string> &names) {
    for (const auto &name : names) m_vector_arrays[name].assign(m_particles.size(), vec_t());
}

std::vector<real> &Simulation::get_scalar_array(const std::string &name) {
    auto it = m_scalar_arrays.find(name);
    if (it == m_scalar_arrays.end()) throw SPHException("no scalar array named " + name);
    return it->second;
}

std::vector<vec_t> &Simulation::get_vector_array(const std::string &name) {
    auto it = m_vector_arrays.find(name);
    if (it == m_vector_arrays.end()) throw SPHException("no vector array named " + name);
    if (!m_host_current) pull();
    m_host_dirty = true;
    return it->second;
}

std::array<double, 7> Simulation::energy() const {
    std::array<double, 7> e{};
    check_status(sphgpu_energy(m_ctx, e.data()), "sphgpu_energy");
    return e;
}

}  // namespace sph
