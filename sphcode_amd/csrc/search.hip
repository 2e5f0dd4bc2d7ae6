// Neighbour search: one wave-coherent tree walk per 64 Morton-consecutive
// particles, writing per-particle candidate lists (BHTree::neighbor_search,
// src/bhtree.cpp:114-126, 234-270).
//
// Gather search (PreInteraction): H = h_guess * kernel_ratio with
// h_guess = (N_ngb m / (rho A_D))^(1/D) (src/pre_interaction.cpp:61-64);
// a cell is pruned when max_k |x_i - c|_k^2 > (edge/2 + H)^2, and a leaf
// member j is a candidate when |x_i - x_j|^2 < H^2.  The candidate's |r_ij|
// is stored beside its index for the Newton iterations.
// Symmetric search (FluidForce): H = max(h_i, largest h in the cell)
// (src/bhtree.cpp:237).
//
// The reference sorts each list by distance (src/bhtree.cpp:120-124) and
// stops at the first r >= h; the pair loops here test r < h on the
// unsorted list instead, which selects the same set.
//
// List layout: entry k of the particle in slot (wave w, lane l) lives at
// (w * cap + k) * 64 + l, so a wave's k-th loads are one contiguous line.
#include "walk.hpp"

namespace sphgpu {

struct SearchArgs {
    int *list;
    double *list_r;      // may be null
    int *cnt;
    int cap;
    int guess;           // 1: set sml = h_guess * ratio first
    double ratio;
    double nngb;
    int64_t *counters;
};

template <int D, bool IJ> struct SearchVisitor {
    const TreeDev &t;
    const BoxDev &box;
    const State &s;
    double xi[D];
    double h, H, H2;
    int count = 0;
    int cap;
    int *out;
    double *outr;

    __device__ __forceinline__ SearchVisitor(const TreeDev &t_, const BoxDev &b_, const State &s_) : t(t_), box(b_), s(s_) {}

    __device__ __forceinline__ bool open(int node)
    {
#pragma clang fp contract(off)
        const double4 g = t.geo[node];
        if constexpr(IJ) {
            const double k = t.ks[node];
            H = h < k ? k : h;
        } else {
            H = h;
        }
        H2 = H * H;
        const double l2 = sq(g.w * 0.5 + H);
        double d[D];
        d[0] = xi[0] - g.x;
        if constexpr(D > 1) d[1] = xi[1] - g.y;
        if constexpr(D > 2) d[2] = xi[2] - g.z;
        min_image<D>(box, d);
        double dmax = d[0] * d[0];
#pragma unroll
        for(int k = 1; k < D; ++k) {
            const double dk = d[k] * d[k];
            dmax = dk > dmax ? dk : dmax;
        }
        return dmax <= l2;
    }

    __device__ __forceinline__ void far(int) {}

    __device__ __forceinline__ void leaf(int4 r, bool open)
    {
        for(int j = r.x; j < r.y; ++j) {
            double xj[D], d[D];
#pragma unroll
            for(int k = 0; k < D; ++k) xj[k] = s.f[F_X + k][j];
            const double r2 = sep2<D>(box, xi, xj, d);
            if(open && r2 < H2) {
                if(count < cap) {
                    out[(size_t)count * WAVE] = j;
                    if(outr) outr[(size_t)count * WAVE] = sqrt(r2);
                }
                ++count;
            }
        }
    }
};

template <int D, bool IJ>
__global__ void __launch_bounds__(BLOCK) search_kernel(State s, TreeDev t, BoxDev box, int n, SearchArgs a)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const bool valid = i < n;
    const size_t wave = (size_t)i / WAVE;
    const int lane = i % WAVE;
    SearchVisitor<D, IJ> v(t, box, s);
    v.cap = a.cap;
    v.out = a.list + wave * a.cap * WAVE + lane;
    v.outr = a.list_r ? a.list_r + wave * a.cap * WAVE + lane : nullptr;
    v.h = 0.0;
    if(valid) {
#pragma unroll
        for(int k = 0; k < D; ++k) v.xi[k] = s.f[F_X + k][i];
        if(a.guess) {
            // p_i.sml = (N_ngb m / (rho A))^(1/D) * ratio (src/pre_interaction.cpp:61-64)
            const double m = s.f[F_MASS][i];
            const double rho = s.f[F_DENS][i];
            v.h = root_dim<D>(a.nngb * m / (rho * UnitVolume<D>::value)) * a.ratio;
            s.f[F_SML][i] = v.h;
        } else {
            v.h = s.f[F_SML][i];
        }
    }
    tree_walk(t, *t.num_nodes, valid, v);
    int over = 0, len = 0;
    if(valid) {
        a.cnt[i] = v.count < a.cap ? v.count : a.cap;
        over = v.count > a.cap ? 1 : 0;
        len = v.count;
    }
    over = wave_sum(over);
    len = wave_maxv(len);
    if(wave_lane() == 0) {
        if(over) atomicAdd((unsigned long long *)&a.counters[C_OVERFLOW], (unsigned long long)over);
        atomicMax((unsigned long long *)&a.counters[C_MAXLIST], (unsigned long long)len);
    }
}

template <int D, bool IJ> static void launch_search(Ctx &c, const SearchArgs &a)
{
    BoxDev box{};
    box.on = c.p.periodic_is_valid;
    for(int k = 0; k < 3; ++k) {
        box.lo[k] = c.p.range_min[k];
        box.hi[k] = c.p.range_max[k];
        box.len[k] = c.p.range_max[k] - c.p.range_min[k];
    }
    const int n = (int)c.n;
    hipLaunchKernelGGL((search_kernel<D, IJ>), dim3(grid_for(n)), dim3(BLOCK), 0, c.stream, c.s, c.t, box, n, a);
}

template <bool IJ> static int search_dispatch(Ctx &c, const SearchArgs &a)
{
    switch(c.D) {
    case 1: launch_search<1, IJ>(c, a); break;
    case 2: launch_search<2, IJ>(c, a); break;
    default: launch_search<3, IJ>(c, a); break;
    }
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

int search_gather(Ctx &c, bool initial)
{
    SearchArgs a{};
    a.list = c.list;
    a.list_r = c.list_r;
    a.cnt = c.list_cnt;
    a.cap = c.cap;
    a.guess = 1;
    a.ratio = initial ? 1.0 : (c.p.iterative_sml ? 1.2 : 1.0);  // src/pre_interaction.cpp:31-35
    a.nngb = (double)c.p.neighbor_number;
    a.counters = c.counters;
    return search_dispatch<false>(c, a);
}

int search_symmetric(Ctx &c)
{
    SearchArgs a{};
    a.list = c.list;
    a.list_r = nullptr;
    a.cnt = c.list_cnt;
    a.cap = c.cap;
    a.guess = 0;
    a.counters = c.counters;
    return search_dispatch<true>(c, a);
}

}  // namespace sphgpu
