#include "sph/snapshot.hpp"

#include <cstdint>
#include <cstring>
#include <fstream>

#include "sph/exception.hpp"

namespace sph {

namespace {

constexpr char MAGIC[8] = {'S', 'P', 'H', 'S', 'N', 'A', 'P', '2'};
constexpr char MAGIC1[8] = {'S', 'P', 'H', 'S', 'N', 'A', 'P', '1'};

template <class T> void put(std::ofstream &f, const T &v) { f.write(reinterpret_cast<const char *>(&v), sizeof(T)); }

template <class T> T get(std::ifstream &f, const std::string &path) {
    T v;
    if (!f.read(reinterpret_cast<char *>(&v), sizeof(T))) throw SPHException(path + ": truncated snapshot");
    return v;
}

// member accessors in file order
using VecField = vec_t SPHParticle::*;
using ScalarField = double SPHParticle::*;
using IntField = int SPHParticle::*;
const VecField kVec[] = {&SPHParticle::pos, &SPHParticle::vel, &SPHParticle::vel_p, &SPHParticle::acc};
const ScalarField kScalar[] = {&SPHParticle::mass,  &SPHParticle::dens,    &SPHParticle::pres,  &SPHParticle::ene,
                               &SPHParticle::ene_p, &SPHParticle::dene,    &SPHParticle::sml,   &SPHParticle::sound,
                               &SPHParticle::balsara, &SPHParticle::alpha, &SPHParticle::gradh, &SPHParticle::phi};
const IntField kInt[] = {&SPHParticle::id, &SPHParticle::neighbor};

}  // namespace

void write_snapshot(const std::string &path, const Snapshot &s) {
    if (s.dim < 1 || s.dim > 3) throw SPHException("write_snapshot: dim must be 1, 2 or 3");
    std::ofstream f(path, std::ios::binary);
    if (!f) throw SPHException(path + ": cannot open for writing");
    f.write(MAGIC, 8);
    put<int32_t>(f, s.dim);
    put<int32_t>(f, 0);
    put<int64_t>(f, (int64_t)s.particles.size());
    put<double>(f, s.time);
    put<double>(f, s.dt);
    put<double>(f, s.h_per_v_sig);
    for (VecField m : kVec)
        for (const auto &p : s.particles)
            for (int k = 0; k < s.dim; ++k) put<double>(f, (p.*m)[k]);
    for (ScalarField m : kScalar)
        for (const auto &p : s.particles) put<double>(f, p.*m);
    for (IntField m : kInt)
        for (const auto &p : s.particles) put<int32_t>(f, (int32_t)(p.*m));
    if (!f) throw SPHException(path + ": write failed");
}

Snapshot read_snapshot(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw SPHException(path + ": cannot open file");
    char magic[8];
    if (!f.read(magic, 8)) throw SPHException(path + ": not a snapshot");
    const bool v2 = std::memcmp(magic, MAGIC, 8) == 0;
    if (!v2 && std::memcmp(magic, MAGIC1, 8) != 0) throw SPHException(path + ": not a snapshot");
    Snapshot s;
    s.dim = get<int32_t>(f, path);
    (void)get<int32_t>(f, path);
    const int64_t n = get<int64_t>(f, path);
    if (s.dim < 1 || s.dim > 3 || n < 0 || n > (int64_t)1 << 31) throw SPHException(path + ": bad snapshot header");
    s.time = get<double>(f, path);
    s.dt = get<double>(f, path);
    if (v2) s.h_per_v_sig = get<double>(f, path);
    s.particles.resize((size_t)n);
    for (VecField m : kVec)
        for (auto &p : s.particles)
            for (int k = 0; k < s.dim; ++k) (p.*m)[k] = get<double>(f, path);
    for (ScalarField m : kScalar)
        for (auto &p : s.particles) p.*m = get<double>(f, path);
    for (IntField m : kInt)
        for (auto &p : s.particles) p.*m = get<int32_t>(f, path);
    return s;
}

}  // namespace sph
