// GravityForce on the device: Barnes-Hut walk per particle with the
// reference's opening test and softening (GravityForce::calculation,
// src/gravity_force.cpp:52-89 -> BHNode::calc_force, src/bhtree.cpp:301-331).
//
// A cell is opened when edge^2 > theta^2 |x_i - com|^2; an opened leaf is
// summed particle by particle, i itself included (the self-potential), with
// the Hernquist & Katz spline softening averaged over h_i and h_j; an
// accepted cell adds its monopole.  phi is reset; the walk's acceleration
// goes to its own buffer and is then added to the hydrodynamic
// acceleration (gravity_combine), so the walk may run while the force pass
// assigns acc.  The walk is the wave-coherent
// pre-order walk of walk.hpp, so each lane keeps its own opening decisions.
//
// T is the interaction arithmetic.  The opening test always runs in FP64,
// so both precisions take the same interaction lists; in mixed mode the
// separations are taken in FP64 relative to particle i, rounded to FP32,
// and the FP32 terms are accumulated in FP64 (Acc in ctx.hpp).
#include "walk.hpp"

namespace sphgpu {

struct GravArgs {
    double G;
    double theta2;
    int64_t *counters;
    const GravNode *gnode;
    const void *gpart;
    double *gacc;        // D x stride, the walk's acceleration
    int64_t stride;
    LeafHit *leafhit;    // split walk records, LEAFHIT_CAP per wave
    int *leafcnt;
};

// acc += gravity's acceleration (after the force pass assigned acc)
template <int D> __global__ void __launch_bounds__(BLOCK) gravity_combine_kernel(State s, int n, const double *gacc, int64_t stride)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
#pragma unroll
    for(int k = 0; k < D; ++k) s.f[F_A + k][i] += gacc[k * stride + i];
}

// Node records: centre of mass, G m and edge^2 from the tree's moments and
// geometry, the particle range, the successor and the leaf flag.
__global__ void __launch_bounds__(BLOCK) pack_nodes_kernel(TreeDev t, GravNode *out, double G, int64_t cap)
{
    const int total = *t.num_nodes;
    for(int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < total && v < cap; v += (int64_t)gridDim.x * BLOCK) {
        const double4 g = t.geo[v];
        const double4 m = t.mom[v];
        const int4 r = t.rng[v];
        GravNode o;
        o.cx = m.x;
        o.cy = m.y;
        o.cz = m.z;
        o.gm = G * m.w;
        o.edge2 = g.w * g.w;
        o.begin = r.x;
        o.end = r.y;
        o.skip = r.z > (int)v ? r.z : (int)v + 1;
        o.leaf = (r.w & 256) != 0;
        o.gmf = (float)o.gm;
        o.pad = 0;
        out[v] = o;
    }
}

// Particle records in tree order: position, mass and 2/h in the pair type
// (FP32 mirrors in mixed mode, as the pair kernels read them).
template <int D, class T> __global__ void __launch_bounds__(BLOCK) pack_parts_kernel(State s, int n, GravPart<T> *out)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i > n) return;
    GravPart<T> p{};
    if(i < n) {
#pragma unroll
        for(int k = 0; k < 3; ++k) p.x[k] = k < D ? s.f[F_X + k][i] : 0.0;
        p.m = fld<T>(s, F_MASS, i);
        p.einv = T(2) * fld<T>(s, F_HINV, i);
    }
    out[i] = p;  // i == n: the spare record
}

// RECORD: the walk of the split pass, which records each opened leaf (with
// the lanes that opened it) instead of summing its particles.
template <int D, class T, bool RECORD = false> struct GravityVisitor {
    const BoxDev &box;
    const GravPart<T> *gp;
    double G, theta2;
    double xi[D];
    T einv_i;   // 2 / h_i
    Acc<T> phi;
    Acc<T> acc[D];
    LeafHit *rec = nullptr;  // RECORD: this wave's records
    int nrec = 0;            // RECORD: leaves opened so far (may pass the capacity)

    __device__ __forceinline__ GravityVisitor(const BoxDev &b_, const GravPart<T> *gp_) : box(b_), gp(gp_) {}

    // Opening test edge^2 > theta^2 |x_i - com|^2 (FP64, all lanes); the
    // lanes of `act` that do not open the node add its monopole, the others
    // add zero.  |x_i - com|^2 is summed left to right without contraction,
    // as the reference's inner product is, so both take the same cells.
    // Returns the lanes of `act` that open the node.
    __device__ __forceinline__ uint64_t node(const GravNode &g, uint64_t act)
    {
#pragma clang fp contract(off)
        double d[D];
        d[0] = xi[0] - g.cx;
        if constexpr(D > 1) d[1] = xi[1] - g.cy;
        if constexpr(D > 2) d[2] = xi[2] - g.cz;
        min_image<D>(box, d);
        double d2 = d[0] * d[0];
#pragma unroll
        for(int k = 1; k < D; ++k) d2 = d2 + d[k] * d[k];
        const uint64_t om = act & ballot(g.edge2 > theta2 * d2);
        const bool far = lane_in(act & ~om);
        T gmn;
        if constexpr(sizeof(T) == 4) gmn = g.gmf;
        else gmn = g.gm;
        const T gm = far ? gmn : T(0);
        const T rinv = rsqrt_t<T>(far ? (T)d2 : T(1));
        phi.add(-(gm * rinv));
        const T c = gm * rinv * rinv * rinv;
#pragma unroll
        for(int k = 0; k < D; ++k) acc[k].add_prod(d[k], -c);
        return om;
    }

    __device__ __forceinline__ void leaf(int b, int e, uint64_t open)
    {
        if constexpr(RECORD) {
            if(nrec < LEAFHIT_CAP && wave_lane() == 0) {
                LeafHit h;
                h.b = b;
                h.e = e;
                h.mask = open;
                rec[nrec] = h;
            }
            ++nrec;
            return;
        }
        const T w = lane_in(open) ? T(1) : T(0);
        for(int j = b; j < e; ++j) pair(load_ro(gp + j), w);
    }

    // The pair (i, j) with weight w: 1, or 0 for a lane that only keeps the
    // wave in step (its sums are then unchanged).
    __device__ __forceinline__ void pair(const GravPart<T> &p, T w)
    {
        double xj[D];
#pragma unroll
        for(int k = 0; k < D; ++k) xj[k] = p.x[k];
        const T mj = p.m;
        const T einv_j = p.einv;
        double dd[D];
        const double r2d = sep2<D>(box, xi, xj, dd);
        T rij[D];
        T r2 = T(0);
#pragma unroll
        for(int k = 0; k < D; ++k) {
            rij[k] = (T)dd[k];
            r2 += rij[k] * rij[k];
        }
        if constexpr(sizeof(T) == 8) r2 = r2d;
        const T rinv = r2 > T(0) ? rsqrt_t<T>(r2) : T(0);
        const T rr = r2 * rinv;
        // far field of both softening lengths unless one is within 2e
        T pot = rinv + rinv, acc3 = T(2) * rinv * rinv * rinv;
        const T ui = rr * einv_i, uj = rr * einv_j;
        if(ui < T(2) || uj < T(2)) {
            T pi_, ai, pj, aj;
            soft_both(ui, einv_i, rinv, pi_, ai);
            soft_both(uj, einv_j, rinv, pj, aj);
            pot = pi_ + pj;
            acc3 = ai + aj;
        }
        const T gm = (T)G * mj * T(0.5) * w;
        phi.add(-(gm * pot));
        const T c = gm * acc3;
#pragma unroll
        for(int k = 0; k < D; ++k) acc[k].add_prod(dd[k], -c);
    }
};

// Split pass, part 1: the walk with monopoles only (no h needed), which
// records the opened leaves per wave for part 2.
template <int D, class T, bool STATS> __global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(GravityOcc<T>::value))) gravity_walk_kernel(State s, TreeDev t, BoxDev box, int n, GravArgs a)
{
    const int i = xcd_block() * BLOCK + threadIdx.x;
    const int task = uniform(i / WAVE);
    const bool valid = i < n;
    GravityVisitor<D, T, true> v(box, nullptr);
    v.G = a.G;
    v.theta2 = a.theta2;
    v.einv_i = T(0);
    v.rec = a.leafhit + (int64_t)task * LEAFHIT_CAP;
#pragma unroll
    for(int k = 0; k < D; ++k) v.xi[k] = valid ? s.f[F_X + k][i] : 0.0;
    __shared__ int stk_node[BLOCK / WAVE][WALK_STACK];
    __shared__ uint64_t stk_mask[BLOCK / WAVE][WALK_STACK];
    const int w = uniform(threadIdx.x / WAVE);
    const WalkCount wc = tree_walk<STATS>(a.gnode, uniform(*t.num_nodes), valid, v, stk_node[w], stk_mask[w]);
    if(valid) {
        s.f[F_PHI][i] = v.phi.val();
#pragma unroll
        for(int k = 0; k < D; ++k) a.gacc[k * a.stride + i] = v.acc[k].val();
    }
    if(wave_lane() == 0) {
        a.leafcnt[task] = v.nrec;
        if(v.nrec > LEAFHIT_CAP) atomicAdd((unsigned long long *)&a.counters[C_GRAV_REWALK], 1ull);
        if constexpr(STATS) {
            stat_max(a.counters, task, C_GRAV_MAXREC, (unsigned)v.nrec);
            stat_add(a.counters, task, C_GRAV_PP, (unsigned)wc.pairs);
            stat_add(a.counters, task, C_GRAV_PN, (unsigned)wc.accepted);
            stat_add(a.counters, task, C_GRAV_OPEN, (unsigned)wc.opened);
            stat_add(a.counters, task, C_GRAV_STEPS, (unsigned)wc.visits);
        }
    }
}

// Split pass, part 2: the particle pairs of the recorded leaves (softened
// with h_i and h_j), added to part 1's sums.  A wave whose record overflowed
// walks the tree again in full and overwrites part 1's sums.
template <int D, class T, bool STATS> __global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(GravityOcc<T>::value))) gravity_leaf_kernel(State s, TreeDev t, BoxDev box, int n, GravArgs a)
{
    const int i = xcd_block() * BLOCK + threadIdx.x;
    const int task = uniform(i / WAVE);
    const bool valid = i < n;
    const int nrec = uniform(a.leafcnt[task]);
    double xi[D];
#pragma unroll
    for(int k = 0; k < D; ++k) xi[k] = valid ? s.f[F_X + k][i] : 0.0;
    const T einv_i = valid ? (T)(2.0 * s.f[F_HINV][i]) : T(0);
    if(nrec > LEAFHIT_CAP) {
        GravityVisitor<D, T> v(box, (const GravPart<T> *)a.gpart);
        v.G = a.G;
        v.theta2 = a.theta2;
        v.einv_i = einv_i;
#pragma unroll
        for(int k = 0; k < D; ++k) v.xi[k] = xi[k];
        __shared__ int stk_node[BLOCK / WAVE][WALK_STACK];
        __shared__ uint64_t stk_mask[BLOCK / WAVE][WALK_STACK];
        const int w = uniform(threadIdx.x / WAVE);
        const WalkCount wc = tree_walk<STATS>(a.gnode, uniform(*t.num_nodes), valid, v, stk_node[w], stk_mask[w]);
        if(STATS && wave_lane() == 0) stat_add(a.counters, task, C_GRAV_SLOTS, (int64_t)WAVE * (unsigned)wc.slots);
        if(valid) {
            s.f[F_PHI][i] = v.phi.val();
#pragma unroll
            for(int k = 0; k < D; ++k) a.gacc[k * a.stride + i] = v.acc[k].val();
        }
        return;
    }
    GravityVisitor<D, T> v(box, (const GravPart<T> *)a.gpart);
    v.G = a.G;
    v.theta2 = a.theta2;
    v.einv_i = einv_i;
#pragma unroll
    for(int k = 0; k < D; ++k) v.xi[k] = xi[k];
    // Each lane runs through its own leaves only: the records' lane masks
    // are transposed into per-lane record masks, 64 records per word, and
    // every lane steps through the (leaf, particle) sequence of its set
    // bits at its own pace, so no lane waits on leaves that it did not
    // open.  A lane adds its pairs in the same order as a sweep over all
    // records would (records in walk order, particles in tree order).  The
    // records are taken LEAFHIT_CHUNK at a time, in order.
    const LeafHit *rec = a.leafhit + (int64_t)task * LEAFHIT_CAP;
    const int lane = wave_lane();
    const int w = uniform(threadIdx.x / WAVE);
    __shared__ int2 srec[BLOCK / WAVE][LEAFHIT_CHUNK];
    constexpr int NW = LEAFHIT_CHUNK / WAVE;
    static_assert(NW == 4, "four record words per lane");
    const GravPart<T> *gp = (const GravPart<T> *)a.gpart;
    int iters = 0;
    for(int base = 0; base < nrec; base += LEAFHIT_CHUNK) {
        const int cnt = min(nrec - base, LEAFHIT_CHUNK);
        const int nwords = (cnt + WAVE - 1) / WAVE;
        uint64_t mine[NW];
#pragma unroll
        for(int c = 0; c < NW; ++c) {
            mine[c] = 0;
            if(c < nwords) {
                const int k = c * WAVE + lane;
                uint64_t m = 0;
                if(k < cnt) {
                    const LeafHit h = rec[base + k];
                    m = h.mask;
                    srec[w][k] = make_int2(h.b, h.e);
                }
                mine[c] = wave_transpose64(m);
            }
        }
        wave_lds_sync();
        int c = 0, j = 0, e = 0;
        uint64_t cur = mine[0];
        bool live = true;
        while(true) {
            // next leaf of this lane once its current one is done
            while(live && j == e) {
                while(cur == 0 && c + 1 < nwords) {
                    ++c;
                    cur = c == 1 ? mine[1] : (c == 2 ? mine[2] : mine[NW - 1]);
                }
                if(cur == 0) {
                    live = false;
                } else {
                    const int r = (int)__builtin_ctzll(cur);
                    cur &= cur - 1;
                    const int2 be = srec[w][c * WAVE + r];
                    j = be.x;
                    e = be.y;
                }
            }
            if(ballot(live) == 0) break;
            // two consecutive particles of the leaf per pass (one if it has
            // one left), so that the leaf bookkeeping above runs half as often
            const bool two = live && j + 1 < e;
            const GravPart<T> p0 = gp[live ? j : 0];
            const GravPart<T> p1 = gp[two ? j + 1 : 0];
            v.pair(p0, live ? T(1) : T(0));
            v.pair(p1, two ? T(1) : T(0));
            j += two ? 2 : (live ? 1 : 0);
            ++iters;
        }
        wave_lds_sync();  // srec is refilled by the next round
    }
    if(STATS && lane == 0) stat_add(a.counters, task, C_GRAV_SLOTS, (int64_t)(2 * WAVE) * (unsigned)iters);
    if(valid) {
        s.f[F_PHI][i] += v.phi.val();
#pragma unroll
        for(int k = 0; k < D; ++k) a.gacc[k * a.stride + i] += v.acc[k].val();
    }
}

static GravArgs grav_args(Ctx &c)
{
    GravArgs a{};
    a.G = c.p.gravity_constant;
    a.theta2 = c.p.gravity_theta * c.p.gravity_theta;
    a.counters = c.counters;
    a.gnode = c.gnode;
    a.gpart = c.gpart;
    a.gacc = c.gacc;
    a.stride = c.cap_n;
    a.leafhit = c.leafhit;
    a.leafcnt = c.leafcnt;
    return a;
}

template <class T, bool STATS> static void walk_d(Ctx &c, const GravArgs &a, const BoxDev &box, hipStream_t st)
{
    const int n = (int)c.n;
    switch(c.D) {
    case 1: hipLaunchKernelGGL((gravity_walk_kernel<1, T, STATS>), dim3(grid_for(n)), dim3(BLOCK), 0, st, c.s, c.t, box, n, a); break;
    case 2: hipLaunchKernelGGL((gravity_walk_kernel<2, T, STATS>), dim3(grid_for(n)), dim3(BLOCK), 0, st, c.s, c.t, box, n, a); break;
    default: hipLaunchKernelGGL((gravity_walk_kernel<3, T, STATS>), dim3(grid_for(n)), dim3(BLOCK), 0, st, c.s, c.t, box, n, a); break;
    }
}

template <class T, bool STATS> static void leaves_d(Ctx &c, const GravArgs &a, const BoxDev &box, hipStream_t st)
{
    const int n = (int)c.n;
    switch(c.D) {
    case 1: hipLaunchKernelGGL((gravity_leaf_kernel<1, T, STATS>), dim3(grid_for(n)), dim3(BLOCK), 0, st, c.s, c.t, box, n, a); break;
    case 2: hipLaunchKernelGGL((gravity_leaf_kernel<2, T, STATS>), dim3(grid_for(n)), dim3(BLOCK), 0, st, c.s, c.t, box, n, a); break;
    default: hipLaunchKernelGGL((gravity_leaf_kernel<3, T, STATS>), dim3(grid_for(n)), dim3(BLOCK), 0, st, c.s, c.t, box, n, a); break;
    }
}

template <class T> static void walk_t(Ctx &c, const GravArgs &a, const BoxDev &box, hipStream_t st)
{
    const int n = (int)c.n;
    hipLaunchKernelGGL(pack_nodes_kernel, dim3(std::min<int64_t>(grid_for(c.node_cap), 2048)), dim3(BLOCK), 0, st, c.t,
                       c.gnode, a.G, c.node_cap);
    if(c.stats_on) walk_d<T, true>(c, a, box, st);
    else walk_d<T, false>(c, a, box, st);
}

template <class T> static void leaves_t(Ctx &c, const GravArgs &a, const BoxDev &box, hipStream_t st)
{
    const int n = (int)c.n;
    // the pre-interaction writes the particle records with h; pack them
    // here only if it has not run on this tree
    GravPart<T> *gp = (GravPart<T> *)c.gpart;
    if(!c.gpart_fresh) {
        switch(c.D) {
        case 1: hipLaunchKernelGGL((pack_parts_kernel<1, T>), dim3(grid_for(n + 1)), dim3(BLOCK), 0, st, c.s, n, gp); break;
        case 2: hipLaunchKernelGGL((pack_parts_kernel<2, T>), dim3(grid_for(n + 1)), dim3(BLOCK), 0, st, c.s, n, gp); break;
        default: hipLaunchKernelGGL((pack_parts_kernel<3, T>), dim3(grid_for(n + 1)), dim3(BLOCK), 0, st, c.s, n, gp); break;
        }
    }
    if(c.stats_on) leaves_d<T, true>(c, a, box, st);
    else leaves_d<T, false>(c, a, box, st);
}

int gravity_walk(Ctx &c, hipStream_t st)
{
    if(!c.p.gravity_is_valid) return SPHGPU_OK;
    if(!c.gnode || !c.gpart || !c.gacc || !c.leafhit) return set_error(SPHGPU_E_STATE, "gravity: no node records");
    const GravArgs a = grav_args(c);
    const BoxDev box = make_box(c);
    if(c.p.precision == SPHGPU_MIXED) walk_t<float>(c, a, box, st);
    else walk_t<double>(c, a, box, st);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

int gravity_leaves(Ctx &c, hipStream_t st)
{
    if(!c.p.gravity_is_valid) return SPHGPU_OK;
    const GravArgs a = grav_args(c);
    const BoxDev box = make_box(c);
    if(c.p.precision == SPHGPU_MIXED) leaves_t<float>(c, a, box, st);
    else leaves_t<double>(c, a, box, st);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

int gravity_combine(Ctx &c, hipStream_t st)
{
    if(!c.p.gravity_is_valid) return SPHGPU_OK;
    const int n = (int)c.n;
    switch(c.D) {
    case 1: hipLaunchKernelGGL(gravity_combine_kernel<1>, dim3(grid_for(n)), dim3(BLOCK), 0, st, c.s, n, c.gacc, c.cap_n); break;
    case 2: hipLaunchKernelGGL(gravity_combine_kernel<2>, dim3(grid_for(n)), dim3(BLOCK), 0, st, c.s, n, c.gacc, c.cap_n); break;
    default: hipLaunchKernelGGL(gravity_combine_kernel<3>, dim3(grid_for(n)), dim3(BLOCK), 0, st, c.s, n, c.gacc, c.cap_n); break;
    }
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

// The stage call (GravityForce::calculation): both parts and the combine
// in order on the context stream, the same arithmetic as sphgpu_integrate.
int gravity_compute(Ctx &c)
{
    int rc = gravity_walk(c, c.stream);
    if(!rc) rc = gravity_leaves(c, c.stream);
    return rc ? rc : gravity_combine(c, c.stream);
}

}  // namespace sphgpu
