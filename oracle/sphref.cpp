/*
 * oracle/sphref.cpp -- CPU restatement of sphcode's per-step SPH particle
 * pipeline (mitchiinaga/sphcode), written from the reference's algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the
 * HIP path and the CPU baseline timed beside it.  Only tests/, the smoke
 * check of __graft_entry__.py and bench.py's cpu_baseline leg load it; the
 * product (sphcode_amd/) never does.
 *
 * It follows the reference step for step, including its data structures
 * where they decide results or cost: an array-of-structures particle
 * vector with intrusive linked lists, a top-down pointer octree built
 * serially, per-particle neighbour lists allocated per pass and sorted by
 * distance, and per-particle OpenMP loops.  Reference citations are
 * file:line in mitchiinaga/sphcode.
 *
 * DIM is a compile-time constant (include/defines.hpp:5); the recipe in
 * oracle/Makefile builds one library per DIM and flag set.  Exported
 * symbols carry the prefix sphref_ and take the same parameter and particle
 * structs as the product's C ABI (include/sphgpu.h).
 *
 * Pinning: every function here is checked by tests/test_oracle_*.py against
 * closed forms, invariants, brute force or special cases; see the header of
 * each test module for which check pins which function.
 */
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/sphgpu.h"

#ifndef SPHREF_DIM
#error "SPHREF_DIM must be defined (1, 2 or 3)"
#endif

namespace sphref {

constexpr int D = SPHREF_DIM;
constexpr int NCHILD = 1 << D;  // include/bhtree.hpp:15

static thread_local std::string g_error;

static int fail(int code, const std::string &msg)
{
    g_error = msg;
    return code;
}

/* ---------------------------------------------------------------- vectors */
// vec_t (include/vector_type.hpp:7-192): fixed-DIM arithmetic, dot products
// summed left to right (include/vector_type.hpp:194-225).
struct Vec {
    double c[D];
    Vec() { for(int k = 0; k < D; ++k) c[k] = 0.0; }
    explicit Vec(double v) { for(int k = 0; k < D; ++k) c[k] = v; }
    double &operator[](int k) { return c[k]; }
    double operator[](int k) const { return c[k]; }
};
static inline Vec operator+(const Vec &a, const Vec &b) { Vec r; for(int k = 0; k < D; ++k) r[k] = a[k] + b[k]; return r; }
static inline Vec operator-(const Vec &a, const Vec &b) { Vec r; for(int k = 0; k < D; ++k) r[k] = a[k] - b[k]; return r; }
static inline Vec operator*(const Vec &a, double s) { Vec r; for(int k = 0; k < D; ++k) r[k] = a[k] * s; return r; }
static inline Vec &operator+=(Vec &a, const Vec &b) { for(int k = 0; k < D; ++k) a[k] += b[k]; return a; }
static inline Vec &operator-=(Vec &a, const Vec &b) { for(int k = 0; k < D; ++k) a[k] -= b[k]; return a; }
static inline double dot(const Vec &a, const Vec &b)
{
    double s = a[0] * b[0];
    for(int k = 1; k < D; ++k) s = s + a[k] * b[k];
    return s;
}
static inline double norm2(const Vec &a) { return dot(a, a); }
static inline double norm(const Vec &a) { return std::sqrt(dot(a, a)); }

static inline double sq(double x) { return x * x; }
static inline double cube(double x) { return x * x * x; }

// h^DIM (include/kernel/kernel_function.hpp:8-16)
static inline double pow_dim(double h)
{
    return D == 1 ? h : D == 2 ? h * h : h * h * h;
}
// h^(DIM-1) (src/pre_interaction.cpp:217-225)
static inline double pow_dim_minus1(double h)
{
    return D == 1 ? 1.0 : D == 2 ? h : h * h;
}
// x^(1/DIM).  The reference writes std::pow(x, 1.0 / DIM)
// (src/pre_interaction.cpp:64); its -ffast-math build (Makefile:7) turns
// that into the identity, sqrt and cbrt for DIM 1, 2, 3.
static inline double root_dim(double x)
{
    return D == 1 ? x : D == 2 ? std::sqrt(x) : std::cbrt(x);
}
// unit-sphere volume A_D (src/pre_interaction.cpp:61-63)
constexpr double kUnitVolume = D == 1 ? 2.0 : D == 2 ? M_PI : 4.0 * M_PI / 3.0;

/* ---------------------------------------------------------------- kernels */
// max(x, 0) written as the reference does, 0.5 * (x + |x|), which is exact.
static inline double pos_part(double x) { return 0.5 * (x + std::fabs(x)); }

// Cubic spline, support radius h (H = h/2, q < 2):
// include/kernel/cubic_spline.hpp:12-52.
struct Cubic {
    static constexpr double sigma = D == 1 ? 2.0 / 3.0 : D == 2 ? 10.0 / (7.0 * M_PI) : 1.0 / M_PI;
    static double w(double r, double h)
    {
        const double hh = h * 0.5;
        const double q = r / hh;
        return sigma / pow_dim(hh) * (0.25 * cube(pos_part(2.0 - q)) - cube(pos_part(1.0 - q)));
    }
    static Vec dw(const Vec &rij, double r, double h)
    {
        if(r == 0.0) return Vec();
        const double hh = h * 0.5;
        const double q = r / hh;
        const double c = -sigma / (pow_dim(hh) * hh * r) * (0.75 * sq(pos_part(2.0 - q)) - 3.0 * sq(pos_part(1.0 - q)));
        return rij * c;
    }
    static double dhw(double r, double h)
    {
        const double hh = h * 0.5;
        const double q = r / hh;
        const double a = 0.5 * (std::fabs(2.0 - q) + 2.0 - q);
        const double b = 0.5 * (std::fabs(1.0 - q) + 1.0 - q);
        return 0.5 * sigma / (pow_dim(hh) * hh) *
               (sq(a) * ((3.0 + D) * 0.25 * q - 0.5 * D) + sq(b) * ((-3.0 - D) * q + D));
    }
};

// Wendland C4, support radius h: include/kernel/wendland_kernel.hpp:10-49
// (sigma = 0 in 1D, where the kernel is not defined).
struct WendlandC4 {
    static constexpr double sigma = D == 1 ? 0.0 : D == 2 ? 9.0 / M_PI : 495.0 / (32.0 * M_PI);
    static double w(double r, double h)
    {
        const double q = r / h;
        const double t = pos_part(1.0 - q);
        return sigma / pow_dim(h) * (t * t * t * t * t * t) * (1.0 + 6.0 * q + 35.0 / 3.0 * q * q);
    }
    static Vec dw(const Vec &rij, double r, double h)
    {
        const double q = r / h;
        const double t = pos_part(1.0 - q);
        const double c = -56.0 / 3.0 * sigma / (pow_dim(h) * sq(h)) * (t * t * t * t * t) * (1.0 + 5.0 * q);
        return rij * c;
    }
    static double dhw(double r, double h)
    {
        const double q = r / h;
        const double t = pos_part(1.0 - q);
        return -sigma / (pow_dim(h) * h * 3.0) * (t * t * t * t * t) *
               (3.0 * D + 15.0 * D * q + (-56.0 + 17.0 * D) * q * q - 35.0 * (8.0 + D) * cube(q));
    }
};

struct Kernel {
    int type;
    double w(double r, double h) const { return type == SPHGPU_WENDLAND ? WendlandC4::w(r, h) : Cubic::w(r, h); }
    Vec dw(const Vec &rij, double r, double h) const { return type == SPHGPU_WENDLAND ? WendlandC4::dw(rij, r, h) : Cubic::dw(rij, r, h); }
    double dhw(double r, double h) const { return type == SPHGPU_WENDLAND ? WendlandC4::dhw(r, h) : Cubic::dhw(r, h); }
};

/* --------------------------------------------------------------- periodic */
// Periodic (include/periodic.hpp:11-72): minimum image chosen among
// r, r+L, r-L per axis with the reference's tie order, and a one-period wrap.
struct Box {
    bool on = false;
    Vec lo, hi, len;
    Vec rij(const Vec &a, const Vec &b) const
    {
        Vec d = a - b;
        if(!on) return d;
        Vec out;
        for(int k = 0; k < D; ++k) {
            const double d1 = d[k];
            const double d2 = d[k] + len[k];
            const double d3 = d[k] - len[k];
            const double a1 = std::fabs(d1), a2 = std::fabs(d2), a3 = std::fabs(d3);
            if(a1 <= a2 && a1 <= a3) out[k] = d1;
            else if(a2 <= a3 && a2 <= a1) out[k] = d2;
            else out[k] = d3;
        }
        return out;
    }
    void wrap(Vec &x) const
    {
        if(!on) return;
        for(int k = 0; k < D; ++k) {
            if(x[k] < lo[k]) x[k] += len[k];
            else if(x[k] > hi[k]) x[k] -= len[k];
        }
    }
};

/* -------------------------------------------------------------- particles */
// SPHParticle (include/particle.hpp:8-33); `next` is an index here.
struct Particle {
    Vec pos, vel, vel_p, acc;
    double mass = 0, dens = 0, pres = 0, ene = 0, ene_p = 0, dene = 0, sml = 0, sound = 0;
    double balsara = 0, alpha = 0, gradh = 0, phi = 0;
    int id = 0, neighbor = 0;
    int next = -1;
};

/* ------------------------------------------------------------------- tree */
// BHTree (include/bhtree.hpp:17-82, src/bhtree.cpp).
struct Node {
    int first;             // head of the particle linked list
    double mass;
    int num;
    int child[NCHILD];     // -1 = absent
    Vec center, com;
    double edge;
    int level;
    double kernel_size;
    bool leaf;
    void reset()
    {
        first = -1; mass = 0.0; num = 0;
        for(int c = 0; c < NCHILD; ++c) child[c] = -1;
        center = Vec(); com = Vec(); edge = 0.0; level = 0; kernel_size = 0.0; leaf = false;
    }
};

struct Tree {
    int max_level = 20, leaf_num = 1;
    bool periodic = false;
    Node root;                 // root lives outside the pool, as in the reference
    std::vector<Node> pool;    // BHTree::resize: particle_num * tree_size nodes
    int used = 0;
    double G = 1.0, theta2 = 0.25;
    bool pool_error = false;

    Node &node(int idx) { return idx < 0 ? root : pool[idx]; }
    const Node &node(int idx) const { return idx < 0 ? root : pool[idx]; }
};

/* ------------------------------------------------------------------ stats */
struct Stats {
    int64_t steps = 0, gather_pairs = 0, force_pairs = 0, gravity_pp = 0, gravity_pn = 0;
    int64_t newton_failures = 0, newton_iterations = 0, search_truncations = 0;
    int64_t list_overflows = 0, max_list_length = 0, tree_nodes = 0, tree_max_level = 0, gravity_opened = 0;
};

/* ------------------------------------------------------------- simulation */
struct Sim {
    sphgpu_params p;
    Kernel kernel;
    Box box;
    Tree tree;
    std::vector<Particle> parts;
    int n = 0;
    double time = 0, dt = 0, h_per_v_sig = 0;
    bool first_pre = true;
    bool exhaustive = false;    // EXHAUSTIVE_SEARCH (include/defines.hpp:32)
    bool tree_ready = false;
    double kernel_ratio = 1.2;
    int capacity = 0;
    // GSPH extra arrays (src/solver.cpp:373-385)
    std::vector<Vec> grad_d, grad_p;
    std::vector<Vec> grad_v[D];
    // recorded sets
    bool record = false;
    std::vector<std::vector<int>> set_gather, set_force;
    Stats stats;
    double stage_seconds[7] = {0, 0, 0, 0, 0, 0, 0};
};

/* ------------------------------------------------------------- tree build */
// BHNode::assign (src/bhtree.cpp:163-204): child slot from strict `>`
// comparisons against the centre (bit k for axis k), child created on first
// use with half the edge and centre offset by a quarter edge; particle
// prepended to the child's list.
static void tree_assign(Tree &t, int parent_idx, std::vector<Particle> &ps, int pi)
{
    Node &par = t.node(parent_idx);
    Particle &p = ps[pi];
    int slot = 0;
    for(int k = 0; k < D; ++k)
        if(p.pos[k] > par.center[k]) slot |= 1 << k;

    int ci = par.child[slot];
    if(ci < 0) {
        if(t.used >= (int)t.pool.size()) {
            // The reference checks `remaind < 0` before taking a node
            // (src/bhtree.cpp:179-181), which lets one node past the pool be
            // written; this restatement stops at the pool size.
            t.pool_error = true;
            return;
        }
        ci = t.used++;
        par.child[slot] = ci;
        Node &c = t.pool[ci];
        c.reset();
        c.level = par.level + 1;
        c.edge = par.edge * 0.5;
        for(int k = 0; k < D; ++k) {
            const double sgn = ((slot >> k) & 1) ? 1.0 : -1.0;
            c.center[k] = par.center[k] + sgn * par.edge * 0.25;
        }
    }
    Node &c = t.pool[ci];
    c.num++;
    c.mass += p.mass;
    c.com += p.pos * p.mass;
    p.next = c.first;
    c.first = pi;
}

// BHNode::create_tree (src/bhtree.cpp:136-161): distribute the node's list
// over its children, then finish each child in slot order: divide the mass
// moment, and split it if it holds more than leaf_num particles and the
// parent level is below max_level; otherwise mark it a leaf.
static void tree_create(Tree &t, int idx, std::vector<Particle> &ps)
{
    {
        Node &nd = t.node(idx);
        for(int c = 0; c < NCHILD; ++c) nd.child[c] = -1;
    }
    int pi = t.node(idx).first;
    while(pi >= 0) {
        const int nxt = ps[pi].next;
        tree_assign(t, idx, ps, pi);
        if(t.pool_error) return;
        pi = nxt;
    }
    for(int c = 0; c < NCHILD; ++c) {
        const int ci = t.node(idx).child[c];
        if(ci < 0) continue;
        Node &ch = t.pool[ci];
        for(int k = 0; k < D; ++k) ch.com[k] /= ch.mass;
        if(ch.num > t.leaf_num && t.node(idx).level < t.max_level) {
            tree_create(t, ci, ps);
            if(t.pool_error) return;
        } else {
            ch.leaf = true;
        }
    }
}

// BHTree::make (src/bhtree.cpp:55-107).
static int tree_make(Sim &s)
{
    Tree &t = s.tree;
    Node &r = t.root;
    // root_clear (include/bhtree.hpp:46-53): mass and centre of mass of the
    // root stay zero; they are never accumulated.
    r.first = -1; r.mass = 0.0; r.num = 0; r.com = Vec(); r.kernel_size = 0.0; r.leaf = false;
    const int n = s.n;
    if(!t.periodic) {
        Vec lo(std::numeric_limits<double>::max()), hi(std::numeric_limits<double>::lowest());
        for(int i = 0; i < n; ++i)
            for(int k = 0; k < D; ++k) {
                lo[k] = std::min(lo[k], s.parts[i].pos[k]);
                hi[k] = std::max(hi[k], s.parts[i].pos[k]);
            }
        r.center = (hi + lo) * 0.5;
        const Vec range = hi - lo;
        double l = 0.0;
        for(int k = 0; k < D; ++k) if(l < range[k]) l = range[k];
        r.edge = l;
    }
    for(int i = 0; i < n - 1; ++i) s.parts[i].next = i + 1;
    s.parts[n - 1].next = -1;
    r.first = 0;
    t.used = 0;
    t.pool_error = false;
    tree_create(t, -1, s.parts);
    if(t.pool_error) return fail(SPHGPU_E_TREE, "There is no free node.");
    s.tree_ready = true;
    s.stats.tree_nodes = t.used + 1;
    int ml = 1;
    for(int i = 0; i < t.used; ++i) ml = std::max(ml, t.pool[i].level);
    s.stats.tree_max_level = ml;
    return SPHGPU_OK;
}

// BHNode::set_kernel (src/bhtree.cpp:206-232): largest h in each subtree.
static double tree_set_kernel(Tree &t, int idx, const std::vector<Particle> &ps)
{
    Node &nd = t.node(idx);
    double kmax = 0.0;
    if(nd.leaf) {
        for(int pi = nd.first; pi >= 0; pi = ps[pi].next)
            if(ps[pi].sml > kmax) kmax = ps[pi].sml;
    } else {
        for(int c = 0; c < NCHILD; ++c) {
            const int ci = nd.child[c];
            if(ci < 0) continue;
            const double h = tree_set_kernel(t, ci, ps);
            if(h > kmax) kmax = h;
        }
    }
    t.node(idx).kernel_size = kmax;
    return kmax;
}

// BHNode::neighbor_search (src/bhtree.cpp:234-270): prune by the L-infinity
// distance from the cell centre against edge/2 + H, H = h_i (gather) or
// max(h_i, cell max-h) (symmetric); collect leaf members with r^2 < H^2.
static void tree_search(const Sim &s, int idx, const Particle &pi, std::vector<int> &list, int &count, bool is_ij, int64_t &overflow)
{
    const Node &nd = s.tree.node(idx);
    const double h = is_ij ? std::max(pi.sml, nd.kernel_size) : pi.sml;
    const double h2 = h * h;
    const double l2 = sq(nd.edge * 0.5 + h);
    const Vec d = s.box.rij(pi.pos, nd.center);
    double dmax = d[0] * d[0];
    for(int k = 1; k < D; ++k) {
        const double dk = d[k] * d[k];
        if(dk > dmax) dmax = dk;
    }
    if(dmax > l2) return;
    if(nd.leaf) {
        for(int pj = nd.first; pj >= 0; pj = s.parts[pj].next) {
            const Vec rij = s.box.rij(pi.pos, s.parts[pj].pos);
            if(norm2(rij) < h2) {
                if(count < (int)list.size()) list[count] = s.parts[pj].id;
                else ++overflow;
                ++count;
            }
        }
    } else {
        for(int c = 0; c < NCHILD; ++c)
            if(nd.child[c] >= 0) tree_search(s, nd.child[c], pi, list, count, is_ij, overflow);
    }
}

// exhaustive_search (src/exhaustive_search.cpp:11-42): r^2 < h_i^2 (gather)
// or r^2 < max(h_i^2, h_j^2) (symmetric) over all particles.
static int brute_search(const Sim &s, const Particle &pi, std::vector<int> &list, bool is_ij, int64_t &overflow)
{
    const double hi2 = pi.sml * pi.sml;
    int count = 0;
    for(int j = 0; j < s.n; ++j) {
        const Particle &pj = s.parts[j];
        const double r2 = norm2(s.box.rij(pi.pos, pj.pos));
        const double lim = is_ij ? std::max(hi2, pj.sml * pj.sml) : hi2;
        if(r2 < lim) {
            if(count < (int)list.size()) list[count] = j;
            else ++overflow;
            ++count;
        }
    }
    return count;
}

// BHTree::neighbor_search (src/bhtree.cpp:114-126) / exhaustive_search:
// collect, then sort by squared distance.
static int neighbor_search(const Sim &s, const Particle &pi, std::vector<int> &list, bool is_ij, int64_t &overflow)
{
    int count = 0;
    if(s.exhaustive) {
        count = brute_search(s, pi, list, is_ij, overflow);
    } else {
        tree_search(s, -1, pi, list, count, is_ij, overflow);
    }
    const int m = std::min(count, (int)list.size());
    const Vec &xi = pi.pos;
    std::sort(list.begin(), list.begin() + m, [&](int a, int b) {
        return norm2(s.box.rij(xi, s.parts[a].pos)) < norm2(s.box.rij(xi, s.parts[b].pos));
    });
    return m;
}

// BHNode::calc_force (src/bhtree.cpp:273-331): Hernquist & Katz spline
// softening with e = h/2, averaged over h_i and h_j.
static inline double soft_pot(double r, double h)
{
    const double e = h * 0.5;
    const double u = r / e;
    if(u < 1.0) return (-0.5 * u * u * (1.0 / 3.0 - 3.0 / 20 * u * u + u * u * u / 20) + 1.4) / e;
    if(u < 2.0) return -1.0 / (15 * r) + (-u * u * (4.0 / 3.0 - u + 0.3 * u * u - u * u * u / 30) + 1.6) / e;
    return 1 / r;
}
static inline double soft_acc(double r, double h)
{
    const double e = h * 0.5;
    const double u = r / e;
    if(u < 1.0) return (4.0 / 3.0 - 1.2 * u * u + 0.5 * u * u * u) / (e * e * e);
    if(u < 2.0) return (-1.0 / 15 + 8.0 / 3 * u * u * u - 3 * u * u * u * u + 1.2 * u * u * u * u * u - u * u * u * u * u * u / 6.0) / (r * r * r);
    return 1 / (r * r * r);
}

// Open a cell when edge^2 > theta^2 |x_i - com|^2; leaves are summed
// particle by particle (i included, which gives the self-potential),
// accepted cells contribute their monopole.
static void tree_gravity(const Sim &s, int idx, Particle &pi, int64_t &npp, int64_t &npn, int64_t &nopen)
{
    const Tree &t = s.tree;
    const Node &nd = s.tree.node(idx);
    const double l2 = nd.edge * nd.edge;
    const Vec d = s.box.rij(pi.pos, nd.com);
    const double d2 = norm2(d);
    const double G = t.G;
    if(l2 > t.theta2 * d2) {
        ++nopen;
        if(nd.leaf) {
            for(int pj = nd.first; pj >= 0; pj = s.parts[pj].next) {
                const Particle &q = s.parts[pj];
                const Vec rij = s.box.rij(pi.pos, q.pos);
                const double r = norm(rij);
                pi.phi -= G * q.mass * (soft_pot(r, pi.sml) + soft_pot(r, q.sml)) * 0.5;
                pi.acc -= rij * (G * q.mass * (soft_acc(r, pi.sml) + soft_acc(r, q.sml)) * 0.5);
                ++npp;
            }
        } else {
            for(int c = 0; c < NCHILD; ++c)
                if(nd.child[c] >= 0) tree_gravity(s, nd.child[c], pi, npp, npn, nopen);
        }
    } else {
        const double rinv = 1.0 / std::sqrt(d2);
        pi.phi -= G * nd.mass * rinv;
        pi.acc -= d * (G * nd.mass * cube(rinv));
        ++npn;
    }
}

/* ------------------------------------------------------- pre-interaction */
// Guess of the smoothing length (src/pre_interaction.cpp:61-64).
static inline double h_guess(const Sim &s, const Particle &p, double ratio)
{
    return root_dim(s.p.neighbor_number * p.mass / (p.dens * kUnitVolume)) * ratio;
}

// PreInteraction::newton_raphson (src/pre_interaction.cpp:227-283) and the
// DISPH variant (src/disph/d_pre_interaction.cpp:174-230): solve
// rho(h) h^D = m N_ngb / A_D (SSPH, GSPH) or n(h) h^D = N_ngb / A_D (DISPH)
// over the pre-fetched sorted list; the iterate after the converging step is
// returned; after 10 iterations the guess is returned.
static double newton_h(Sim &s, const Particle &pi, const std::vector<int> &list, int cnt, bool number_density,
                       int64_t &iters, int64_t &failures)
{
    double h = pi.sml / s.kernel_ratio;
    const double b = number_density ? s.p.neighbor_number / kUnitVolume : pi.mass * s.p.neighbor_number / kUnitVolume;
    const double eps = 1e-4;
    for(int it = 0; it < 10; ++it) {
        ++iters;
        const double hb = h;
        double rho = 0.0, drho = 0.0;
        for(int n = 0; n < cnt; ++n) {
            const Particle &pj = s.parts[list[n]];
            const double r = norm(s.box.rij(pi.pos, pj.pos));
            if(r >= h) break;
            const double m = number_density ? 1.0 : pj.mass;
            if(number_density) {
                rho += s.kernel.w(r, h);
                drho += s.kernel.dhw(r, h);
            } else {
                rho += m * s.kernel.w(r, h);
                drho += m * s.kernel.dhw(r, h);
            }
        }
        const double f = rho * pow_dim(h) - b;
        const double df = drho * pow_dim(h) + D * rho * pow_dim_minus1(h);
        h -= f / df;
        if(std::fabs(h - hb) < (h + hb) * eps) return h;
    }
    ++failures;
    return pi.sml / s.kernel_ratio;
}

// PreInteraction::initial_smoothing (src/pre_interaction.cpp:171-215):
// first call only; h from the guess without the kernel ratio, density from
// the initial-condition density guess.
static void initial_smoothing(Sim &s)
{
    const int n = s.n;
    const int cap = s.capacity;
    int64_t overflow = 0;
#pragma omp parallel for reduction(+ : overflow)
    for(int i = 0; i < n; ++i) {
        Particle &pi = s.parts[i];
        std::vector<int> list(cap);
        pi.sml = h_guess(s, pi, 1.0);
        const int cnt = neighbor_search(s, pi, list, false, overflow);
        double rho = 0.0;
        for(int k = 0; k < cnt; ++k) {
            const Particle &pj = s.parts[list[k]];
            const double r = norm(s.box.rij(pi.pos, pj.pos));
            if(r >= pi.sml) break;
            rho += pj.mass * s.kernel.w(r, pi.sml);
        }
        pi.dens = rho;
    }
    s.stats.list_overflows += overflow;
}

// 2D curl is a scalar, 3D a vector (include/vector_type.hpp:239-268).
static inline double comp(const Vec &v, int k) { return k < D ? v.c[k < D ? k : 0] : 0.0; }
struct Curl {
    double c[3] = {0, 0, 0};  // 2D uses c[2] only
    void add(const Vec &a, const Vec &b, double m)
    {
        if(D == 3) {
            c[0] += (comp(a, 1) * comp(b, 2) - comp(a, 2) * comp(b, 1)) * m;
            c[1] += (comp(a, 2) * comp(b, 0) - comp(a, 0) * comp(b, 2)) * m;
        }
        c[2] += (comp(a, 0) * comp(b, 1) - comp(a, 1) * comp(b, 0)) * m;
    }
    void mul(double f) { for(double &x : c) x *= f; }
    void div(double f) { for(double &x : c) x /= f; }
    double magnitude() const
    {
        if(D == 2) return std::fabs(c[2]);
        return std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    }
};

// PreInteraction::calculation for the three formulations:
//   SSPH  src/pre_interaction.cpp:39-169
//   DISPH src/disph/d_pre_interaction.cpp:21-162
//   GSPH  src/gsph/g_pre_interaction.cpp:27-144
static int pre_interaction(Sim &s)
{
    if(s.first_pre) {
        initial_smoothing(s);
        s.first_pre = false;
    }
    const int n = s.n;
    const int type = s.p.sph_type;
    const double gamma = s.p.gamma;
    const bool balsara = s.p.av_use_balsara_switch && D != 1;
    const bool tdav = s.p.av_use_time_dependent;
    const double amax = s.p.av_alpha_max, amin = s.p.av_alpha_min, eps_av = s.p.av_epsilon;
    const double dt = s.dt;
    const int cap = s.capacity;
    const bool second = type == SPHGPU_GSPH && s.p.gsph_is_2nd_order;
    if(s.record) s.set_gather.assign(n, {});

    double hvs_min = std::numeric_limits<double>::max();
    int64_t iters = 0, failures = 0, overflow = 0, gpairs = 0, trunc = 0, maxlen = 0;
#pragma omp parallel for reduction(min : hvs_min) reduction(+ : iters, failures, overflow, gpairs, trunc) reduction(max : maxlen)
    for(int i = 0; i < n; ++i) {
        Particle &pi = s.parts[i];
        std::vector<int> list(cap);
        pi.sml = h_guess(s, pi, s.kernel_ratio);
        const double search_h = pi.sml;
        const int ncand = neighbor_search(s, pi, list, false, overflow);
        if(ncand > maxlen) maxlen = ncand;
        if(s.record) {
            std::vector<int> sorted(list.begin(), list.begin() + ncand);
            std::sort(sorted.begin(), sorted.end());
            s.set_gather[i] = sorted;
        }
        if(s.p.iterative_sml) {
            pi.sml = newton_h(s, pi, list, ncand, type == SPHGPU_DISPH, iters, failures);
            if(pi.sml > search_h) ++trunc;
        }

        const Vec &xi = pi.pos;
        const double hi = pi.sml;
        double rho = 0.0, drho = 0.0;        // SSPH / GSPH
        double pbar = 0.0, dpbar = 0.0, nd = 0.0, dnd = 0.0;  // DISPH
        double vsig_max = pi.sound * 2.0;
        int count = 0;
        for(int k = 0; k < ncand; ++k) {
            const int j = list[k];
            const Particle &pj = s.parts[j];
            const Vec rij = s.box.rij(xi, pj.pos);
            const double r = norm(rij);
            if(r >= hi) break;
            ++count;
            const double wij = s.kernel.w(r, hi);
            if(type == SPHGPU_DISPH) {
                const double dhw = s.kernel.dhw(r, hi);
                rho += pj.mass * wij;
                nd += wij;
                pbar += pj.mass * pj.ene * wij;
                dpbar += pj.mass * pj.ene * dhw;
                dnd += dhw;
            } else {
                rho += pj.mass * wij;
                if(type == SPHGPU_SSPH) drho += pj.mass * s.kernel.dhw(r, hi);
            }
            if(i != j) {
                const double vsig = pi.sound + pj.sound - 3.0 * dot(rij, pi.vel - pj.vel) / r;
                if(vsig > vsig_max) vsig_max = vsig;
            }
        }
        gpairs += count;
        pi.dens = rho;
        if(type == SPHGPU_DISPH) {
            pi.pres = (gamma - 1.0) * pbar;
            // f_ij = 1 - gradh_i / (m_j u_j) (src/disph/d_pre_interaction.cpp:97-98)
            pi.gradh = hi / (D * nd) * dpbar / (1.0 + hi / (D * nd) * dnd);
        } else {
            pi.pres = (gamma - 1.0) * rho * pi.ene;
            if(type == SPHGPU_SSPH) pi.gradh = 1.0 / (1.0 + hi / (D * rho) * drho);
        }
        pi.neighbor = count;
        const double hvs = hi / vsig_max;
        if(hvs < hvs_min) hvs_min = hvs;

        if(type == SPHGPU_GSPH) {
            // MUSCL gradients (src/gsph/g_pre_interaction.cpp:113-136)
            if(!second) continue;
            Vec gd, gu, gv[D];
            for(int k = 0; k < count; ++k) {
                const Particle &pj = s.parts[list[k]];
                const Vec rij = s.box.rij(xi, pj.pos);
                const double r = norm(rij);
                const Vec dwij = s.kernel.dw(rij, r, hi);
                gd += dwij * pj.mass;
                gu += dwij * (pj.mass * (pj.ene - pi.ene));
                for(int c = 0; c < D; ++c) gv[c] += dwij * (pj.mass * (pj.vel[c] - pi.vel[c]));
            }
            s.grad_d[i] = gd;
            s.grad_p[i] = (gd * pi.ene + gu) * (gamma - 1.0);
            const double rinv = 1.0 / pi.dens;
            for(int c = 0; c < D; ++c) s.grad_v[c][i] = gv[c] * rinv;
            continue;
        }

        // Artificial-viscosity switches (src/pre_interaction.cpp:115-161,
        // src/disph/d_pre_interaction.cpp:115-153).  DISPH weights the
        // velocity derivatives by m_j u_j and normalises by P_i/(gamma-1).
        if(balsara) {
            double divv = 0.0;
            Curl rot;
            for(int k = 0; k < count; ++k) {
                const Particle &pj = s.parts[list[k]];
                const Vec rij = s.box.rij(xi, pj.pos);
                const double r = norm(rij);
                const Vec dwij = s.kernel.dw(rij, r, hi);
                const Vec vij = pi.vel - pj.vel;
                const double wgt = type == SPHGPU_DISPH ? pj.mass * pj.ene : pj.mass;
                divv -= wgt * dot(vij, dwij);
                rot.add(vij, dwij, wgt);
            }
            if(type == SPHGPU_DISPH) {
                const double p_inv = (gamma - 1.0) / pi.pres;
                divv *= p_inv;
                rot.mul(p_inv);
            } else {
                divv /= pi.dens;
                rot.div(pi.dens);
            }
            pi.balsara = std::fabs(divv) / (std::fabs(divv) + rot.magnitude() + 1e-4 * pi.sound / hi);
            if(tdav) {
                const double tau_inv = eps_av * pi.sound / hi;
                const double da = (-(pi.alpha - amin) * tau_inv + std::max(-divv, 0.0) * (amax - pi.alpha)) * dt;
                pi.alpha += da;
            }
        } else if(tdav) {
            double divv = 0.0;
            for(int k = 0; k < count; ++k) {
                const Particle &pj = s.parts[list[k]];
                const Vec rij = s.box.rij(xi, pj.pos);
                const double r = norm(rij);
                const Vec dwij = s.kernel.dw(rij, r, hi);
                const Vec vij = pi.vel - pj.vel;
                const double wgt = type == SPHGPU_DISPH ? pj.mass * pj.ene : pj.mass;
                divv -= wgt * dot(vij, dwij);
            }
            if(type == SPHGPU_DISPH) divv *= (gamma - 1.0) / pi.pres;
            else divv /= pi.dens;
            const double tau_inv = eps_av * pi.sound / hi;
            const double si = std::max(-divv, 0.0);
            pi.alpha = (pi.alpha + dt * tau_inv * amin + si * dt * amax) / (1.0 + dt * tau_inv + si * dt);
        }
    }
    s.h_per_v_sig = hvs_min;
    s.stats.newton_iterations += iters;
    s.stats.newton_failures += failures;
    s.stats.list_overflows += overflow;
    s.stats.gather_pairs += gpairs;
    s.stats.search_truncations += trunc;
    s.stats.max_list_length = std::max<int64_t>(s.stats.max_list_length, maxlen);
    if(!s.exhaustive) tree_set_kernel(s.tree, -1, s.parts);
    if(overflow) return fail(SPHGPU_E_LIST_OVERFLOW, "neighbour list capacity exceeded");
    return SPHGPU_OK;
}

/* ----------------------------------------------------------- fluid force */
// Monaghan (1997) signal-velocity viscosity (src/fluid_force.cpp:89-106).
static inline double art_visc(const Particle &a, const Particle &b, const Vec &rij)
{
    const Vec vij = a.vel - b.vel;
    const double vr = dot(vij, rij);
    if(!(vr < 0)) return 0.0;
    const double alpha = 0.5 * (a.alpha + b.alpha);
    const double bal = 0.5 * (a.balsara + b.balsara);
    const double w = vr / norm(rij);
    const double vsig = a.sound + b.sound - 3.0 * w;
    const double rho_inv = 2.0 / (a.dens + b.dens);
    return -0.5 * bal * alpha * vsig * w * rho_inv;
}

// Artificial conductivity (src/fluid_force.cpp:108-116).
static inline double art_cond(const Sim &s, const Particle &a, const Particle &b, const Vec &rij, const Vec &dwij)
{
    const double vsig = s.p.gravity_is_valid ? std::fabs(dot(a.vel - b.vel, rij) / norm(rij))
                                             : std::sqrt(2.0 * std::fabs(a.pres - b.pres) / (a.dens + b.dens));
    return s.p.ac_alpha * b.mass * vsig * (a.ene - b.ene) * dot(dwij, rij) / norm(rij);
}

// van Leer limiter (src/gsph/g_fluid_force.cpp:27-36).
static inline double van_leer(double a, double b)
{
    const double ab = a * b;
    return ab <= 0 ? 0.0 : 2.0 * ab / (a + b);
}

// HLL solver with Roe-averaged wave speeds (src/gsph/g_fluid_force.cpp:168-199).
// States are (velocity along e, density, pressure, sound speed).
static inline void hll(const double L[4], const double R[4], double &pstar, double &vstar)
{
    const double sl = std::sqrt(L[1]), sr = std::sqrt(R[1]);
    const double inv = 1.0 / (sl + sr);
    const double ut = (sl * L[0] + sr * R[0]) * inv;
    const double ct = (sl * L[3] + sr * R[3]) * inv;
    const double s_l = std::min(L[0] - L[3], ut - ct);
    const double s_r = std::max(R[0] + R[3], ut + ct);
    const double c1 = L[1] * (s_l - L[0]);
    const double c2 = R[1] * (s_r - R[0]);
    const double c3 = 1.0 / (c1 - c2);
    const double c4 = L[2] - L[0] * c1;
    const double c5 = R[2] - R[0] * c2;
    vstar = (c5 - c4) * c3;
    pstar = (c1 * c5 - c2 * c4) * c3;
}

// FluidForce::calculation:
//   SSPH  src/fluid_force.cpp:26-86
//   DISPH src/disph/d_fluid_force.cpp:26-86
//   GSPH  src/gsph/g_fluid_force.cpp:39-166
static int fluid_force(Sim &s)
{
    const int n = s.n;
    const int type = s.p.sph_type;
    const int cap = s.capacity;
    const double gamma = s.p.gamma;
    const bool use_ac = s.p.ac_is_valid;
    const double dt = s.dt;
    const bool second = s.p.gsph_is_2nd_order;
    if(s.record) s.set_force.assign(n, {});
    int64_t overflow = 0, fpairs = 0, maxlen = 0;
#pragma omp parallel for reduction(+ : overflow, fpairs) reduction(max : maxlen)
    for(int i = 0; i < n; ++i) {
        Particle &pi = s.parts[i];
        std::vector<int> list(cap);
        const int cnt = neighbor_search(s, pi, list, true, overflow);
        if(cnt > maxlen) maxlen = cnt;
        const Vec &xi = pi.pos;
        const Vec &vi = pi.vel;
        const double hi = pi.sml;
        Vec acc;
        double dene = 0.0;
        std::vector<int> members;

        // per-formulation constants of particle i
        const double p_rho2_i = pi.pres / sq(pi.dens);
        const double g2u_i = sq(gamma - 1.0) * pi.ene;
        const double g2u_p_i = g2u_i / pi.pres;
        const double mu_inv_i = 1.0 / (pi.mass * pi.ene);
        const double rho2_inv_i = 1.0 / sq(pi.dens);

        for(int k = 0; k < cnt; ++k) {
            const int j = list[k];
            const Particle &pj = s.parts[j];
            const Vec rij = s.box.rij(xi, pj.pos);
            const double r = norm(rij);
            if(r >= std::max(hi, pj.sml) || r == 0.0) continue;
            ++fpairs;
            if(s.record) members.push_back(j);

            const Vec dwi = s.kernel.dw(rij, r, hi);
            const Vec dwj = s.kernel.dw(rij, r, pj.sml);
            if(type == SPHGPU_GSPH) {
                const double rinv = 1.0 / r;
                const Vec e = rij * rinv;
                const double ve_i = dot(vi, e);
                const double ve_j = dot(pj.vel, e);
                double pstar, vstar;
                if(second) {
                    // MUSCL reconstruction (src/gsph/g_fluid_force.cpp:98-135)
                    double R[4], L[4];
                    const double di = 0.5 * (1.0 - pi.sound * dt * rinv);
                    const double dj = 0.5 * (1.0 - pj.sound * dt * rinv);
                    const double dv = ve_i - ve_j;
                    Vec gvi, gvj;
                    for(int c = 0; c < D; ++c) {
                        gvi[c] = dot(s.grad_v[c][i], e);
                        gvj[c] = dot(s.grad_v[c][j], e);
                    }
                    const double dve_i = dot(gvi, e) * r;
                    const double dve_j = dot(gvj, e) * r;
                    R[0] = ve_i - van_leer(dv, dve_i) * di;
                    L[0] = ve_j + van_leer(dv, dve_j) * dj;
                    const double dd = pi.dens - pj.dens;
                    const double dd_i = dot(s.grad_d[i], e) * r;
                    const double dd_j = dot(s.grad_d[j], e) * r;
                    R[1] = pi.dens - van_leer(dd, dd_i) * di;
                    L[1] = pj.dens + van_leer(dd, dd_j) * dj;
                    const double dp = pi.pres - pj.pres;
                    const double dp_i = dot(s.grad_p[i], e) * r;
                    const double dp_j = dot(s.grad_p[j], e) * r;
                    R[2] = pi.pres - van_leer(dp, dp_i) * di;
                    L[2] = pj.pres + van_leer(dp, dp_j) * dj;
                    R[3] = std::sqrt(gamma * R[2] / R[1]);
                    L[3] = std::sqrt(gamma * L[2] / L[1]);
                    hll(L, R, pstar, vstar);
                } else {
                    const double R[4] = {ve_i, pi.dens, pi.pres, pi.sound};
                    const double L[4] = {ve_j, pj.dens, pj.pres, pj.sound};
                    hll(L, R, pstar, vstar);
                }
                const Vec vst = e * vstar;
                const double rho2_inv_j = 1.0 / sq(pj.dens);
                const Vec f = dwi * (pj.mass * pstar * rho2_inv_i) + dwj * (pj.mass * pstar * rho2_inv_j);
                acc -= f;
                dene -= dot(f, vst - vi);
                continue;
            }

            const Vec dwij = (dwi + dwj) * 0.5;
            const Vec vij = vi - pj.vel;
            const double pij = art_visc(pi, pj, rij);
            const double ac = use_ac ? art_cond(s, pi, pj, rij, dwij) : 0.0;
            if(type == SPHGPU_DISPH) {
                const double fij = 1.0 - pi.gradh / (pj.mass * pj.ene);
                const double fji = 1.0 - pj.gradh * mu_inv_i;
                const double u_p_j = pj.ene / pj.pres;
                acc -= dwi * (pj.mass * (g2u_p_i * pj.ene * fij + 0.5 * pij)) + dwj * (pj.mass * (g2u_i * u_p_j * fji + 0.5 * pij));
                dene += pj.mass * g2u_p_i * pj.ene * fij * dot(vij, dwi) + 0.5 * pj.mass * pij * dot(vij, dwij) + ac;
            } else {
                acc -= dwi * (pj.mass * (p_rho2_i * pi.gradh + 0.5 * pij)) + dwj * (pj.mass * (pj.pres / sq(pj.dens) * pj.gradh + 0.5 * pij));
                dene += pj.mass * p_rho2_i * pi.gradh * dot(vij, dwi) + 0.5 * pj.mass * pij * dot(vij, dwij) + ac;
            }
        }
        pi.acc = acc;
        pi.dene = dene;
        if(s.record) {
            std::sort(members.begin(), members.end());
            s.set_force[i] = members;
        }
    }
    s.stats.list_overflows += overflow;
    s.stats.force_pairs += fpairs;
    s.stats.max_list_length = std::max<int64_t>(s.stats.max_list_length, maxlen);
    if(overflow) return fail(SPHGPU_E_LIST_OVERFLOW, "neighbour list capacity exceeded");
    return SPHGPU_OK;
}

/* ---------------------------------------------------------------- gravity */
// GravityForce::calculation (src/gravity_force.cpp:52-89).
static int gravity_force(Sim &s)
{
    if(!s.p.gravity_is_valid) return SPHGPU_OK;
    const int n = s.n;
    const double G = s.p.gravity_constant;
    int64_t npp = 0, npn = 0, nopen = 0;
#pragma omp parallel for reduction(+ : npp, npn, nopen)
    for(int i = 0; i < n; ++i) {
        Particle &pi = s.parts[i];
        if(s.exhaustive) {
            double phi = 0.0;
            Vec f;
            for(int j = 0; j < n; ++j) {
                const Particle &pj = s.parts[j];
                const Vec rij = s.box.rij(pi.pos, pj.pos);
                const double r = norm(rij);
                phi -= G * pj.mass * (soft_pot(r, pi.sml) + soft_pot(r, pj.sml)) * 0.5;
                f -= rij * (G * pj.mass * (soft_acc(r, pi.sml) + soft_acc(r, pj.sml)) * 0.5);
            }
            pi.acc += f;
            pi.phi = phi;
            npp += n;
        } else {
            // BHTree::tree_force (src/bhtree.cpp:128-132)
            pi.phi = 0.0;
            tree_gravity(s, -1, pi, npp, npn, nopen);
        }
    }
    s.stats.gravity_pp += npp;
    s.stats.gravity_pn += npn;
    s.stats.gravity_opened += nopen;
    return SPHGPU_OK;
}

/* -------------------------------------------------------------- timestep */
// TimeStep::calculation (src/timestep.cpp:18-38).
static double timestep(Sim &s)
{
    double dtf = std::numeric_limits<double>::max();
    const double cf = s.p.cfl_force;
#pragma omp parallel for reduction(min : dtf)
    for(int i = 0; i < s.n; ++i) {
        const double a = norm(s.parts[i].acc);
        if(a > 0.0) {
            const double d = cf * std::sqrt(s.parts[i].sml / a);
            if(d < dtf) dtf = d;
        }
    }
    const double dts = s.p.cfl_sound * s.h_per_v_sig;
    s.dt = std::min(dts, dtf);
    return s.dt;
}

/* ----------------------------------------------------------- kick-drift */
// Solver::predict (src/solver.cpp:431-456).
static void predict(Sim &s)
{
    const double dt = s.dt;
    const double cs = s.p.gamma * (s.p.gamma - 1.0);
#pragma omp parallel for
    for(int i = 0; i < s.n; ++i) {
        Particle &p = s.parts[i];
        p.vel_p = p.vel + p.acc * (0.5 * dt);
        p.ene_p = p.ene + p.dene * (0.5 * dt);
        p.pos += p.vel_p * dt;
        p.vel += p.acc * dt;
        p.ene += p.dene * dt;
        p.sound = std::sqrt(cs * p.ene);
        s.box.wrap(p.pos);
    }
}

// Solver::correct (src/solver.cpp:458-474).
static void correct(Sim &s)
{
    const double dt = s.dt;
    const double cs = s.p.gamma * (s.p.gamma - 1.0);
#pragma omp parallel for
    for(int i = 0; i < s.n; ++i) {
        Particle &p = s.parts[i];
        p.vel = p.vel_p + p.acc * (0.5 * dt);
        p.ene = p.ene_p + p.dene * (0.5 * dt);
        p.sound = std::sqrt(cs * p.ene);
    }
}

using Clock = std::chrono::steady_clock;
static double since(Clock::time_point t0)
{
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

}  // namespace sphref

using namespace sphref;

/* ================================================================ C API == */
extern "C" {

const char *sphref_last_error(void) { return g_error.c_str(); }
int sphref_dim(void) { return D; }

int sphref_create(const sphgpu_params *prm, void **out)
{
    if(!prm || !out) return fail(SPHGPU_E_INVALID, "null argument");
    if(prm->dim != D) return fail(SPHGPU_E_INVALID, "params.dim does not match this oracle build");
    if(prm->kernel != SPHGPU_CUBIC_SPLINE && prm->kernel != SPHGPU_WENDLAND)
        return fail(SPHGPU_E_INVALID, "kernel is unknown.");  // src/simulation.cpp:18-19
    if(prm->sph_type < 0 || prm->sph_type > 2) return fail(SPHGPU_E_INVALID, "Unknown SPH type");
    Sim *s = new Sim();
    s->p = *prm;
    s->kernel.type = prm->kernel;
    s->box.on = prm->periodic_is_valid != 0;
    for(int k = 0; k < D; ++k) {
        s->box.lo[k] = prm->range_min[k];
        s->box.hi[k] = prm->range_max[k];
    }
    s->box.len = s->box.hi - s->box.lo;
    // BHTree::initialize (src/bhtree.cpp:12-40)
    Tree &t = s->tree;
    t.max_level = prm->tree_max_level;
    t.leaf_num = prm->tree_leaf_particle_num;
    t.root.reset();
    t.root.level = 1;
    t.periodic = s->box.on;
    if(t.periodic) {
        t.root.center = (s->box.hi + s->box.lo) * 0.5;
        double l = 0.0;
        for(int k = 0; k < D; ++k) if(l < s->box.len[k]) l = s->box.len[k];
        t.root.edge = l;
    }
    t.G = prm->gravity_constant;
    t.theta2 = prm->gravity_theta * prm->gravity_theta;
    s->kernel_ratio = prm->iterative_sml ? 1.2 : 1.0;  // src/pre_interaction.cpp:31-35
    s->capacity = prm->neighbor_capacity > 0 ? prm->neighbor_capacity : prm->neighbor_number * 20;
    s->time = prm->time_start;
    s->dt = 0.0;
    *out = s;
    return SPHGPU_OK;
}

int sphref_destroy(void *ctx)
{
    delete static_cast<Sim *>(ctx);
    return SPHGPU_OK;
}

int sphref_set_threads(void *, int threads)
{
#ifdef _OPENMP
    if(threads > 0) omp_set_num_threads(threads);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int sphref_set_exhaustive(void *ctx, int on)
{
    static_cast<Sim *>(ctx)->exhaustive = on != 0;
    return SPHGPU_OK;
}

int sphref_set_record_sets(void *ctx, int on)
{
    static_cast<Sim *>(ctx)->record = on != 0;
    return SPHGPU_OK;
}

int sphref_upload(void *ctx, const sphgpu_particles *h)
{
    Sim &s = *static_cast<Sim *>(ctx);
    if(!h || h->n <= 0) return fail(SPHGPU_E_INVALID, "empty particle array");
    const int n = (int)h->n;
    s.n = n;
    s.parts.assign(n, Particle());
    auto getv = [&](const double *a, int i, Vec &v) { if(a) for(int k = 0; k < D; ++k) v[k] = a[i * D + k]; };
    auto gets = [&](const double *a, int i, double &v) { if(a) v = a[i]; };
    for(int i = 0; i < n; ++i) {
        Particle &p = s.parts[i];
        getv(h->pos, i, p.pos); getv(h->vel, i, p.vel); getv(h->vel_p, i, p.vel_p); getv(h->acc, i, p.acc);
        gets(h->mass, i, p.mass); gets(h->dens, i, p.dens); gets(h->pres, i, p.pres); gets(h->ene, i, p.ene);
        gets(h->ene_p, i, p.ene_p); gets(h->dene, i, p.dene); gets(h->sml, i, p.sml); gets(h->sound, i, p.sound);
        gets(h->balsara, i, p.balsara); gets(h->alpha, i, p.alpha); gets(h->gradh, i, p.gradh); gets(h->phi, i, p.phi);
        p.id = i;
        p.neighbor = h->neighbor ? h->neighbor[i] : 0;
    }
    s.grad_d.assign(n, Vec());
    s.grad_p.assign(n, Vec());
    for(int c = 0; c < D; ++c) s.grad_v[c].assign(n, Vec());
    if(h->grad_density) for(int i = 0; i < n; ++i) for(int k = 0; k < D; ++k) s.grad_d[i][k] = h->grad_density[i * D + k];
    if(h->grad_pressure) for(int i = 0; i < n; ++i) for(int k = 0; k < D; ++k) s.grad_p[i][k] = h->grad_pressure[i * D + k];
    if(h->grad_velocity)
        for(int i = 0; i < n; ++i) for(int c = 0; c < D; ++c) for(int k = 0; k < D; ++k)
            s.grad_v[c][i][k] = h->grad_velocity[(i * D + c) * D + k];
    // BHTree::resize (src/bhtree.cpp:42-53): 5 nodes per particle.
    s.tree.pool.assign((size_t)n * 5, Node());
    s.first_pre = true;
    s.tree_ready = false;
    return SPHGPU_OK;
}

int sphref_set_tree_pool(void *ctx, int64_t nodes)
{
    Sim &s = *static_cast<Sim *>(ctx);
    s.tree.pool.assign((size_t)std::max<int64_t>(nodes, 0), Node());
    return SPHGPU_OK;
}

int sphref_download(void *ctx, sphgpu_particles *h)
{
    Sim &s = *static_cast<Sim *>(ctx);
    if(!h || h->n != s.n) return fail(SPHGPU_E_INVALID, "size mismatch");
    auto putv = [&](double *a, int i, const Vec &v) { if(a) for(int k = 0; k < D; ++k) a[i * D + k] = v[k]; };
    auto puts = [&](double *a, int i, double v) { if(a) a[i] = v; };
    for(int i = 0; i < s.n; ++i) {
        const Particle &p = s.parts[i];
        putv(h->pos, i, p.pos); putv(h->vel, i, p.vel); putv(h->vel_p, i, p.vel_p); putv(h->acc, i, p.acc);
        puts(h->mass, i, p.mass); puts(h->dens, i, p.dens); puts(h->pres, i, p.pres); puts(h->ene, i, p.ene);
        puts(h->ene_p, i, p.ene_p); puts(h->dene, i, p.dene); puts(h->sml, i, p.sml); puts(h->sound, i, p.sound);
        puts(h->balsara, i, p.balsara); puts(h->alpha, i, p.alpha); puts(h->gradh, i, p.gradh); puts(h->phi, i, p.phi);
        if(h->id) h->id[i] = p.id;
        if(h->neighbor) h->neighbor[i] = p.neighbor;
        if(h->grad_density) putv(h->grad_density, i, s.grad_d[i]);
        if(h->grad_pressure) putv(h->grad_pressure, i, s.grad_p[i]);
        if(h->grad_velocity) for(int c = 0; c < D; ++c) for(int k = 0; k < D; ++k)
            h->grad_velocity[(i * D + c) * D + k] = s.grad_v[c][i][k];
    }
    return SPHGPU_OK;
}

int sphref_make_tree(void *ctx)
{
    Sim &s = *static_cast<Sim *>(ctx);
    if(s.n <= 0) return fail(SPHGPU_E_STATE, "no particles");
    if(s.exhaustive) return SPHGPU_OK;  // src/solver.cpp:406-410, 422-424
    auto t0 = Clock::now();
    const int rc = tree_make(s);
    s.stage_seconds[2] += since(t0);
    return rc;
}

int sphref_pre_interaction(void *ctx)
{
    Sim &s = *static_cast<Sim *>(ctx);
    if(!s.tree_ready && !s.exhaustive) return fail(SPHGPU_E_STATE, "tree not built");
    auto t0 = Clock::now();
    const int rc = pre_interaction(s);
    s.stage_seconds[3] += since(t0);
    return rc;
}

int sphref_fluid_force(void *ctx)
{
    Sim &s = *static_cast<Sim *>(ctx);
    if(!s.tree_ready && !s.exhaustive) return fail(SPHGPU_E_STATE, "tree not built");
    auto t0 = Clock::now();
    const int rc = fluid_force(s);
    s.stage_seconds[4] += since(t0);
    return rc;
}

int sphref_gravity_force(void *ctx)
{
    Sim &s = *static_cast<Sim *>(ctx);
    if(!s.tree_ready && !s.exhaustive) return fail(SPHGPU_E_STATE, "tree not built");
    auto t0 = Clock::now();
    const int rc = gravity_force(s);
    s.stage_seconds[5] += since(t0);
    return rc;
}

int sphref_timestep(void *ctx, double *dt_out)
{
    Sim &s = *static_cast<Sim *>(ctx);
    auto t0 = Clock::now();
    const double dt = timestep(s);
    s.stage_seconds[0] += since(t0);
    if(dt_out) *dt_out = dt;
    return SPHGPU_OK;
}

int sphref_predict(void *ctx)
{
    Sim &s = *static_cast<Sim *>(ctx);
    auto t0 = Clock::now();
    predict(s);
    s.stage_seconds[1] += since(t0);
    return SPHGPU_OK;
}

int sphref_correct(void *ctx)
{
    Sim &s = *static_cast<Sim *>(ctx);
    auto t0 = Clock::now();
    correct(s);
    s.stage_seconds[6] += since(t0);
    return SPHGPU_OK;
}

// Solver::initialize after the initial condition (src/solver.cpp:392-414).
int sphref_initialize(void *ctx)
{
    Sim &s = *static_cast<Sim *>(ctx);
    if(s.n <= 0) return fail(SPHGPU_E_STATE, "no particles");
    const double cs = s.p.gamma * (s.p.gamma - 1.0);
    for(int i = 0; i < s.n; ++i) {
        s.parts[i].alpha = s.p.av_alpha;
        s.parts[i].balsara = 1.0;
        s.parts[i].sound = std::sqrt(cs * s.parts[i].ene);
    }
    int rc;
    if((rc = sphref_make_tree(ctx))) return rc;
    if((rc = sphref_pre_interaction(ctx))) return rc;
    if((rc = sphref_fluid_force(ctx))) return rc;
    return sphref_gravity_force(ctx);
}

// Solver::integrate (src/solver.cpp:417-429) + Simulation::update_time.
int sphref_integrate(void *ctx, double fixed_dt)
{
    Sim &s = *static_cast<Sim *>(ctx);
    int rc;
    if(fixed_dt > 0) s.dt = fixed_dt;
    else if((rc = sphref_timestep(ctx, nullptr))) return rc;
    if((rc = sphref_predict(ctx))) return rc;
    if((rc = sphref_make_tree(ctx))) return rc;
    if((rc = sphref_pre_interaction(ctx))) return rc;
    if((rc = sphref_fluid_force(ctx))) return rc;
    if((rc = sphref_gravity_force(ctx))) return rc;
    if((rc = sphref_correct(ctx))) return rc;
    s.time += s.dt;
    s.stats.steps++;
    return SPHGPU_OK;
}

int sphref_get_time(void *ctx, double *time, double *dt, double *hvs)
{
    Sim &s = *static_cast<Sim *>(ctx);
    if(time) *time = s.time;
    if(dt) *dt = s.dt;
    if(hvs) *hvs = s.h_per_v_sig;
    return SPHGPU_OK;
}

int sphref_set_time(void *ctx, double time, double dt)
{
    Sim &s = *static_cast<Sim *>(ctx);
    s.time = time;
    s.dt = dt;
    return SPHGPU_OK;
}

// Output::output_energy sums (src/output.cpp:66-90) plus total momentum.
int sphref_energy(void *ctx, double *out)
{
    Sim &s = *static_cast<Sim *>(ctx);
    double ek = 0, et = 0, ep = 0;
    double mom[3] = {0, 0, 0};
    for(int i = 0; i < s.n; ++i) {
        const Particle &p = s.parts[i];
        ek += 0.5 * p.mass * norm2(p.vel);
        et += p.mass * p.ene;
        ep += 0.5 * p.mass * p.phi;
        for(int k = 0; k < D; ++k) mom[k] += p.mass * p.vel[k];
    }
    out[0] = ek; out[1] = et; out[2] = ep; out[3] = ek + et + ep;
    for(int k = 0; k < D; ++k) out[4 + k] = mom[k];
    return SPHGPU_OK;
}

int sphref_get_stats(void *ctx, sphgpu_stats *st)
{
    Sim &s = *static_cast<Sim *>(ctx);
    st->steps = s.stats.steps; st->gather_pairs = s.stats.gather_pairs; st->force_pairs = s.stats.force_pairs;
    st->gravity_pp = s.stats.gravity_pp; st->gravity_pn = s.stats.gravity_pn;
    st->newton_failures = s.stats.newton_failures; st->newton_iterations = s.stats.newton_iterations;
    st->search_truncations = s.stats.search_truncations; st->list_overflows = s.stats.list_overflows;
    st->max_list_length = s.stats.max_list_length; st->tree_nodes = s.stats.tree_nodes;
    st->tree_max_level = s.stats.tree_max_level;
    st->gravity_opened = s.stats.gravity_opened;
    return SPHGPU_OK;
}

int sphref_reset_stats(void *ctx)
{
    Sim &s = *static_cast<Sim *>(ctx);
    s.stats = Stats();
    for(double &x : s.stage_seconds) x = 0.0;
    return SPHGPU_OK;
}

// Wall seconds per stage: timestep, predict, tree, pre, force, gravity, correct.
int sphref_stage_seconds(void *ctx, double *out7)
{
    Sim &s = *static_cast<Sim *>(ctx);
    for(int k = 0; k < 7; ++k) out7[k] = s.stage_seconds[k];
    return SPHGPU_OK;
}

int sphref_neighbor_sets(void *ctx, int kind, int32_t *counts, int32_t *lists, int64_t cap)
{
    Sim &s = *static_cast<Sim *>(ctx);
    const auto &sets = kind == 0 ? s.set_gather : s.set_force;
    if((int)sets.size() != s.n) return fail(SPHGPU_E_STATE, "sets not recorded");
    int64_t off = 0;
    for(int i = 0; i < s.n; ++i) {
        counts[i] = (int32_t)sets[i].size();
        for(int j : sets[i]) {
            if(off >= cap) return fail(SPHGPU_E_INVALID, "list buffer too small");
            lists[off++] = j;
        }
    }
    return SPHGPU_OK;
}

// ---- building blocks, exported for the oracle's own tests ----------------
// W(r, h), grad W = r_ij * c (evaluated with r_ij = (r, 0, ...)), dW/dh.
int sphref_kernel_eval(int kernel, double r, double h, double *w, double *c, double *dhw)
{
    Kernel k{kernel};
    Vec rij;
    rij[0] = r;
    *w = k.w(r, h);
    const Vec g = k.dw(rij, r, h);
    *c = r != 0.0 ? g[0] / r : 0.0;
    *dhw = k.dhw(r, h);
    return SPHGPU_OK;
}

// Hernquist-Katz softened potential kernel f and force kernel g.
int sphref_softening(double r, double h, double *f, double *g)
{
    *f = soft_pot(r, h);
    *g = soft_acc(r, h);
    return SPHGPU_OK;
}

int sphref_hll(const double *L, const double *R, double *pstar, double *vstar)
{
    hll(L, R, *pstar, *vstar);
    return SPHGPU_OK;
}

double sphref_van_leer(double a, double b) { return van_leer(a, b); }

int sphref_min_image(void *ctx, const double *a, const double *b, double *out)
{
    Sim &s = *static_cast<Sim *>(ctx);
    Vec va, vb;
    for(int k = 0; k < D; ++k) { va[k] = a[k]; vb[k] = b[k]; }
    const Vec d = s.box.rij(va, vb);
    for(int k = 0; k < D; ++k) out[k] = d[k];
    return SPHGPU_OK;
}

// Tree nodes in depth-first pre-order (children in slot order).
int sphref_tree_info(void *ctx, int64_t *num_nodes, int32_t *level, int32_t *count, double *edge,
                     double *center, double *mass, double *com, int64_t cap)
{
    Sim &s = *static_cast<Sim *>(ctx);
    if(!s.tree_ready) return fail(SPHGPU_E_STATE, "tree not built");
    int64_t k = 0;
    std::vector<int> stack{-1};
    // root's count is the particle number (the reference never stores it)
    while(!stack.empty()) {
        const int idx = stack.back();
        stack.pop_back();
        const Node &nd = s.tree.node(idx);
        if(k < cap) {
            if(level) level[k] = nd.level;
            if(count) count[k] = idx < 0 ? s.n : nd.num;
            if(edge) edge[k] = nd.edge;
            for(int d = 0; d < D; ++d) {
                if(center) center[k * D + d] = nd.center[d];
                if(com) com[k * D + d] = nd.com[d];
            }
            if(mass) mass[k] = nd.mass;
        }
        ++k;
        if(!nd.leaf)
            for(int c = NCHILD - 1; c >= 0; --c)
                if(nd.child[c] >= 0) stack.push_back(nd.child[c]);
    }
    *num_nodes = k;
    return SPHGPU_OK;
}

}  // extern "C"
