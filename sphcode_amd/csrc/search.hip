// Neighbour search (BHTree::neighbor_search, src/bhtree.cpp:114-126,
// 234-270) for 64 Morton-consecutive particles per wave, in two phases:
//
//  1. Group walk.  The wave walks the pre-order tree once, 64 consecutive
//     nodes at a time (one per lane, visited ones found by a prefix maximum
//     of subtree ends), keeping every leaf whose cube can hold a
//     candidate of any of its particles: per axis, the minimum-image
//     distance between the group's box centre and the cell centre must not
//     exceed group half-extent + edge/2 + H, with H the group's largest
//     search radius (symmetric search: max(that, cell's largest h)).  This
//     is implied by the reference's per-particle test, so no candidate of
//     any lane is lost.  Leaves are collected into LDS.
//  2. Candidate tiles.  The collected leaves' particles (contiguous ranges
//     in key order) are staged 64 at a time into LDS with coalesced loads;
//     every lane then applies the reference's leaf test to each staged
//     particle: |x_i - x_j|^2 < H_i^2 with H_i = h_guess * kernel_ratio
//     (gather, src/pre_interaction.cpp:61-70) or max(h_i, leaf's largest h)
//     (symmetric, src/bhtree.cpp:237, 256), into a 64-bit hit mask, and
//     appends its hits in staging order.
//
// The sets are the reference's; the reference then sorts each list by
// distance (src/bhtree.cpp:120-124) and stops at the first r >= h, whereas
// the pair loops here test r < h on the unsorted list -- the same set.
//
// The combined search (split mode) walks with sml = the gather radius and
// the cells' largest gather radius, keeps pairs with r < max(sml_i, sml_j),
// and files those with r^2 < sml_i^2 (the gather candidates) at the front of
// the list and the rest at the back, so that one walk serves both the
// pre-interaction and the force stage.
//
// List layout: list_at (ctx.hpp).
#include "ctx.hpp"

namespace sphgpu {

struct SearchArgs {
    int *list;
    double *list_r;      // |r_ij| of each gather candidate (FP64 mode)
    float *list_rf;      // the same in FP32 (mixed mode)
    int *cnt;            // entries at the front of the list
    int *cnt2;           // entries at the back of the list
    int split;           // symmetric search that also marks the gather set:
                         // r < sml_i at the front, the rest at the back
    int cap;
    int guess;           // 1: set sml = h_guess * ratio first
    double ratio;
    double nngb;
    int64_t *counters;
    int *work;           // per-XCD task counters, zeroed before the launch
};

constexpr int SW = BLOCK / WAVE;  // waves per block
constexpr int LEAFCAP = 128;      // leaves collected per flush
constexpr int FLATCAP = 256;      // listed (lane, candidate) hits per tile

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src)
{
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, WAVE);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, WAVE);
    return ((uint64_t)hi << 32) | lo;
}

// Wave-wide exclusive prefix sum.
__device__ __forceinline__ int wave_exclusive_scan(int v, int &total)
{
    const int lane = wave_lane();
    int x = v;
#pragma unroll
    for(int off = 1; off < WAVE; off <<= 1) {
        const int y = __shfl_up(x, off, WAVE);
        if(lane >= off) x += y;
    }
    total = __shfl(x, WAVE - 1, WAVE);
    return x - v;
}

// Minimum-image separation of one axis (include/periodic.hpp:34-59).
__device__ __forceinline__ double min_image1(const BoxDev &b, int k, double d)
{
    if(!b.on) return d;
    const double d2 = d + b.len[k], d3 = d - b.len[k];
    const double a1 = fabs(d), a2 = fabs(d2), a3 = fabs(d3);
    double r = d3;
    if(a2 <= a3 && a2 <= a1) r = d2;
    if(a1 <= a2 && a1 <= a3) r = d;
    return r;
}

template <int D, bool IJ>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(OCC_SEARCH))) search_kernel(State s, TreeDev t, BoxDev box, const int2 *groups, const int *num_groups,
                                                      SearchArgs a)
{
    __shared__ int l_begin[SW][LEAFCAP];
    __shared__ int l_pre[SW][LEAFCAP + 1];
    __shared__ double l_ks[SW][LEAFCAP];
    __shared__ double c_x[SW][3][WAVE];
    __shared__ double c_lim[SW][WAVE];  // IJ: the candidate's radius squared
    __shared__ int c_j[SW][WAVE];
    __shared__ uint16_t flat[SW][FLATCAP];

    const int w = uniform(threadIdx.x / WAVE);
    const int lane = wave_lane();
    const int ngroups = uniform(*num_groups);
    // persistent waves: groups are taken one at a time from this XCD's
    // contiguous share of the Morton curve
    TaskShare share(a.work, ngroups);
    for(int grp = share.next(); grp >= 0; grp = share.next()) {
    const int2 gr = make_int2(uniform(groups[grp].x), uniform(groups[grp].y));
    const int i = gr.x + lane;
    const bool valid = i < gr.y;

    double xi[D], h = 0.0;
#pragma unroll
    for(int k = 0; k < D; ++k) xi[k] = valid ? s.f[F_X + k][i] : 0.0;
    if(valid) {
        if(a.guess) {
            // p_i.sml = (N_ngb m / (rho A))^(1/D) * ratio (src/pre_interaction.cpp:61-64)
            const double m = s.f[F_MASS][i];
            const double rho = s.f[F_DENS][i];
            h = root_dim<D>(a.nngb * m / (rho * UnitVolume<D>::value)) * a.ratio;
            s.f[F_SML][i] = h;
        } else {
            h = s.f[F_SML][i];
        }
    }
    const double h2_own = h * h;
    const int64_t h2_bits = __double_as_longlong(h2_own);
    const size_t lbase = list_base(valid ? i : 0, a.cap);

    // group box: centre, half extent, largest radius (wave-uniform)
    double gc[D], ge[D];
#pragma unroll
    for(int k = 0; k < D; ++k) {
        const double lo = wave_minv(valid ? xi[k] : 1.7976931348623157e308);
        const double hi = wave_maxv(valid ? xi[k] : -1.7976931348623157e308);
        gc[k] = uniform_d(0.5 * (lo + hi));
        ge[k] = uniform_d(0.5 * (hi - lo) * (1.0 + 1e-12) + 1e-300);
    }
    const double hmax = uniform_d(wave_maxv(valid ? h : 0.0));

    int *out = a.list;
    double *outr = a.list_r;
    float *outrf = a.list_rf;
    int count = 0, count2 = 0;  // candidates found (front / back)
    int wf = 0, wb = 0;          // of those, written (at most cap in total)

    const int total = uniform(*t.num_nodes);
    int node = 0;
    int64_t visited = 0, tested = 0;
    while(node < total) {
        // ---- phase 1: collect leaves, 64 consecutive pre-order nodes per
        // window.  A node is visited iff every ancestor of it is opened
        // (overlapping and internal).  Within a window the window's first
        // node is visited, and a later node v is visited iff no earlier
        // node of the window that is not opened has its subtree span
        // [u, skip(u)) over v: v >= the exclusive prefix maximum of skip
        // over the window's not-opened nodes.  The next window starts at
        // the larger of the window's end and that maximum at the last lane.
        // Leaves are kept in pre-order, as the one-node-at-a-time walk
        // would keep them.
        int nleaf = 0;
        while(node < total && nleaf <= LEAFCAP - WAVE) {
            node = uniform(node);
            const int v = node + lane;
            const bool in = v < total;
            int closed_skip = 0;
            bool leafhit = false;
            int4 r = make_int4(0, 0, 0, 0);
            double ks = 0.0;
            if(in) {
                const double4 g = t.geo[v];
                r = t.rng[v];
                double H = hmax;
                if constexpr(IJ) {
                    ks = t.ks[v];
                    H = ks > H ? ks : H;
                }
                const double reach = g.w * 0.5 + H;
                const double gcen[3] = {g.x, g.y, g.z};
                bool overlap = true;
#pragma unroll
                for(int k = 0; k < D; ++k) {
                    const double dd = min_image1(box, k, gc[k] - gcen[k]);
                    overlap = overlap && fabs(dd) <= (ge[k] + reach) * (1.0 + 1e-12);
                }
                const bool leaf = (r.w & 256) != 0;
                const int skip = r.z > v ? r.z : v + 1;
                closed_skip = overlap && !leaf ? 0 : skip;
                leafhit = overlap && leaf;
            }
            const int incl = wave_incl_max_nonneg(closed_skip);
            const int excl = wave_shr1(incl);
            const bool vis = in && v >= excl;
            visited += __popcll(__ballot(vis));
            const bool hit = vis && leafhit;
            const uint64_t hits = __ballot(hit);
            if(hit) {
                const int at = nleaf + __popcll(hits & ((1ull << lane) - 1ull));
                l_begin[w][at] = r.x;
                l_pre[w][at] = r.y - r.x;
                l_ks[w][at] = ks;
            }
            nleaf += uniform(__popcll(hits));
            const int far = __builtin_amdgcn_readlane(incl, WAVE - 1);
            node = node + WAVE > far ? node + WAVE : far;
        }
        wave_lds_sync();
        // prefix of leaf sizes (each lane owns two slots)
        int c0 = 0, c1 = 0;
        if(2 * lane < nleaf) c0 = l_pre[w][2 * lane];
        if(2 * lane + 1 < nleaf) c1 = l_pre[w][2 * lane + 1];
        int ncand = 0;
        const int ex = wave_exclusive_scan(c0 + c1, ncand);
        ncand = uniform(ncand);
        wave_lds_sync();
        if(2 * lane < nleaf) l_pre[w][2 * lane] = ex;
        if(2 * lane + 1 < nleaf) l_pre[w][2 * lane + 1] = ex + c0;
        if(lane == 0) l_pre[w][nleaf] = ncand;
        wave_lds_sync();

        // ---- phase 2: candidate tiles of 64
        int lc = 0;  // last leaf starting at or before tb
        for(int tb = 0; tb < ncand; tb += WAVE) {
            // stage: each lane fetches one candidate and keeps it only if it
            // lies within the group's largest radius of the group box (the
            // box contains every x_i, so each lane's own test implies this)
            const int f = tb + lane;
            // the candidate's leaf: the last leaf starting at or before f,
            // found among the few leaves that start inside the tile
            int lo = lc, nx = lc;
            for(int L = lc + 1; L < nleaf; ++L) {
                const int st = uniform(l_pre[w][L]);
                if(st > tb + WAVE) break;
                nx = L;
                lo = f >= st ? L : lo;
            }
            lc = nx;
            bool keep = false;
            int j = 0;
            double xj[D], ksj = 0.0;
            if(f < ncand) {
                j = l_begin[w][lo] + (f - l_pre[w][lo]);
                double d2b = 0.0;
#pragma unroll
                for(int k = 0; k < D; ++k) {
                    xj[k] = s.f[F_X + k][j];
                    const double ex = fabs(min_image1(box, k, xj[k] - gc[k])) - ge[k];
                    d2b += ex > 0.0 ? ex * ex : 0.0;
                }
                double Hc = hmax;
                if constexpr(IJ) {
                    ksj = l_ks[w][lo];
                    Hc = ksj > Hc ? ksj : Hc;
                }
                keep = d2b <= Hc * Hc * (1.0 + 1e-9);
            }
            const uint64_t kept = __ballot(keep);
            if(keep) {
                const int pos = __popcll(kept & ((1ull << lane) - 1ull));
                c_j[w][pos] = j;
#pragma unroll
                for(int k = 0; k < D; ++k) c_x[w][k][pos] = xj[k];
                if constexpr(IJ) {
                    // split mode: r < max(sml_i, sml_j) (each radius bounds
                    // that particle's final h unless truncated); otherwise
                    // the reference's r < max(h_i, leaf's largest h).  As
                    // squares: max(a, b)^2 == max(a^2, b^2) for a, b >= 0.
                    const double kl = a.split ? s.f[F_SML][j] : ksj;
                    c_lim[w][pos] = kl * kl;
                }
            }
            const int m = uniform(__popcll(kept));
            tested += m;
            wave_lds_sync();
            // pass A: every lane tests every staged candidate (uniform
            // loop, broadcast LDS reads) into a hit mask, and a mask of the
            // hits inside its own radius (the gather entries of a split
            // search; every hit otherwise)
            uint64_t mine = 0, mfront = 0;
            if(valid) {
                uint32_t mlo = 0, mhi = 0, flo = 0, fhi = 0;
                auto test = [&](int c, uint32_t &hb, uint32_t &fb) {
                    double xj[D], d[D];
#pragma unroll
                    for(int k = 0; k < D; ++k) xj[k] = c_x[w][k][c];
                    const double r2 = sep2<D>(box, xi, xj, d);
                    // r2 < max(h_i^2, kl^2) as two tests; all three are
                    // non-negative, so their bit patterns order as integers
                    const int64_t rb = __double_as_longlong(r2);
                    const bool f = rb < h2_bits;
                    bool h = f;
                    if constexpr(IJ) h = f | (rb < __double_as_longlong(c_lim[w][c]));
                    hb = (uint32_t)h;
                    fb = (uint32_t)f;
                };
                const int m1 = m < 32 ? m : 32;
                for(int c = 0; c < m1; ++c) {
                    uint32_t hb, fb;
                    test(c, hb, fb);
                    mlo |= hb << c;
                    flo |= fb << c;
                }
                for(int c = 32; c < m; ++c) {
                    uint32_t hb, fb;
                    test(c, hb, fb);
                    mhi |= hb << (c - 32);
                    fhi |= fb << (c - 32);
                }
                mine = ((uint64_t)mhi << 32) | mlo;
                mfront = a.split ? ((uint64_t)fhi << 32) | flo : mine;
            }
            // pass B: append each lane's hits in staging order, gather
            // entries from the front of its list and force-only entries
            // from the back.  The tile's (lane, candidate) hits are listed
            // in LDS and written 64 at a time, one per lane, each at the
            // slot its rank among its lane's hits gives; a tile in which
            // some list would reach its capacity (or with more than FLATCAP
            // hits) takes the one-lane-at-a-time loop, which keeps the
            // first cap entries.
            const int nh = __popcll(mine);
            const int nf = __popcll(mfront);
            int htotal = 0;
            const int hbase = wave_exclusive_scan(nh, htotal);
            htotal = uniform(htotal);
            const bool fits = wf + wb + nh <= a.cap;
            if(uniform64(__ballot(!fits)) == 0 && htotal <= FLATCAP) {
                {
                    uint64_t mm = mine;
                    int k = hbase;
                    while(mm) {
                        const int c = __builtin_ctzll(mm);
                        mm &= mm - 1ull;
                        flat[w][k++] = (uint16_t)((lane << 8) | c);
                    }
                }
                wave_lds_sync();
                for(int r = 0; r < htotal; r += WAVE) {
                    const int p = r + lane;
                    const bool act = p < htotal;
                    const int code = act ? (int)flat[w][p] : 0;
                    const int owner = code >> 8, c = code & 255;
                    const uint64_t om = shfl64(mine, owner), of = shfl64(mfront, owner);
                    const int owf = __shfl(wf, owner, WAVE), owb = __shfl(wb, owner, WAVE);
                    double xo[D];
#pragma unroll
                    for(int k = 0; k < D; ++k) xo[k] = __longlong_as_double((long long)shfl64((uint64_t)__double_as_longlong(xi[k]), owner));
                    if(act) {
                        const uint64_t below = (1ull << c) - 1ull;
                        const int jj = c_j[w][c];
                        const int io = gr.x + owner;
                        const size_t ob = list_base(io, a.cap);
                        if((of >> c) & 1ull) {
                            const size_t at = ob + list_off(owf + __popcll(of & below));
                            out[at] = jj;
                            if(outr || outrf) {
                                double xj[D], d[D];
#pragma unroll
                                for(int k = 0; k < D; ++k) xj[k] = c_x[w][k][c];
                                const double r2 = sep2<D>(box, xo, xj, d);
                                if(outr) outr[at] = sqrt(r2);
                                if(outrf) outrf[at] = gather_r_f32(r2);
                            }
                        } else {
                            out[ob + list_off(a.cap - 1 - (owb + __popcll(om & ~of & below)))] = jj;
                        }
                    }
                }
                wf += nf;
                wb += nh - nf;
                count += nf;
                count2 += nh - nf;
            } else if(valid) {
                uint64_t mm = mine;
                while(mm) {
                    const int c = __builtin_ctzll(mm);
                    mm &= mm - 1ull;
                    double xj[D], d[D];
#pragma unroll
                    for(int k = 0; k < D; ++k) xj[k] = c_x[w][k][c];
                    const double r2 = sep2<D>(box, xi, xj, d);
                    const int jj = c_j[w][c];
                    const bool room = wf + wb < a.cap;
                    if(!a.split || r2 < h2_own) {
                        if(room) {
                            const size_t at = lbase + list_off(wf);
                            out[at] = jj;
                            if(outr) outr[at] = sqrt(r2);
                            if(outrf) outrf[at] = gather_r_f32(r2);
                            ++wf;
                        }
                        ++count;
                    } else {
                        if(room) {
                            out[lbase + list_off(a.cap - 1 - wb)] = jj;
                            ++wb;
                        }
                        ++count2;
                    }
                }
            }
            wave_lds_sync();
        }
    }

    int over = 0, len = 0;
    if(valid) {
        // past the capacity the first cap entries found are kept and the
        // overflow is reported
        a.cnt[i] = wf;
        a.cnt2[i] = wb;
        over = count + count2 > a.cap ? 1 : 0;
        len = count + count2;
    }
    over = wave_sum(over);
    len = wave_maxv(len);
    if(lane == 0) {
        atomicAdd((unsigned long long *)&a.counters[C_SEARCH_NODES], (unsigned long long)visited);
        atomicAdd((unsigned long long *)&a.counters[C_SEARCH_CAND], (unsigned long long)tested);
        if(over) atomicAdd((unsigned long long *)&a.counters[C_OVERFLOW], (unsigned long long)over);
        atomicMax((unsigned long long *)&a.counters[C_MAXLIST], (unsigned long long)len);
    }
    }  // groups
}

template <int D, bool IJ> static void launch_search(Ctx &c, const SearchArgs &a)
{
    const BoxDev box = make_box(c);
    // enough blocks to fill every CU (at most 8 blocks of 4 waves each), a
    // multiple of the XCD count; blocks beyond that find no task and leave
    const int blocks = std::max(NUM_XCD, std::min(grid_for(c.n, WAVE) / SW, 8 * c.num_cu) / NUM_XCD * NUM_XCD);
    reset_work(c, 0);
    SearchArgs aw = a;
    aw.work = c.work;
    hipLaunchKernelGGL((search_kernel<D, IJ>), dim3(blocks), dim3(BLOCK), 0, c.stream, c.s, c.t, box, c.groups, c.num_groups, aw);
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

// sml = h_guess * ratio (src/pre_interaction.cpp:61-64), the gather
// search radius, before the combined search
template <int D> __global__ void __launch_bounds__(BLOCK) guess_kernel(State s, int n, double nngb, double ratio)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
    s.f[F_SML][i] = root_dim<D>(nngb * s.f[F_MASS][i] / (s.f[F_DENS][i] * UnitVolume<D>::value)) * ratio;
}

int search_guess(Ctx &c)
{
    const double ratio = c.p.iterative_sml ? 1.2 : 1.0;
    const double nngb = (double)c.p.neighbor_number;
    switch(c.D) {
    case 1: hipLaunchKernelGGL(guess_kernel<1>, dim3(grid_for(c.n)), dim3(BLOCK), 0, c.stream, c.s, (int)c.n, nngb, ratio); break;
    case 2: hipLaunchKernelGGL(guess_kernel<2>, dim3(grid_for(c.n)), dim3(BLOCK), 0, c.stream, c.s, (int)c.n, nngb, ratio); break;
    default: hipLaunchKernelGGL(guess_kernel<3>, dim3(grid_for(c.n)), dim3(BLOCK), 0, c.stream, c.s, (int)c.n, nngb, ratio); break;
    }
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

// One walk for both sets: with sml = h_guess * ratio and the cells' largest
// sml from tree_set_kernel, the test r < max(sml_i, sml_j) finds every
// gather candidate (r < sml_i, kept at the front with |r_ij|) and every pair
// with r < max(h_i, h_j) for the final h, as long as no final h exceeds its
// search radius (the caller checks the truncation counter).
int search_combined(Ctx &c)
{
    SearchArgs a{};
    a.list = c.list;
    a.list_r = c.list_r;
    a.list_rf = c.list_rf;
    a.cnt = c.list_cnt;
    a.cnt2 = c.list_cnt2;
    a.split = 1;
    a.cap = c.cap;
    a.guess = 0;
    a.counters = c.counters;
    return search_dispatch<true>(c, a);
}

int search_gather(Ctx &c, bool initial)
{
    SearchArgs a{};
    a.list = c.list;
    a.list_r = c.list_r;
    a.list_rf = c.list_rf;
    a.cnt = c.list_cnt;
    a.cnt2 = c.list_cnt2;
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
    a.list_rf = nullptr;
    a.cnt = c.list_cnt;
    a.cnt2 = c.list_cnt2;
    a.cap = c.cap;
    a.guess = 0;
    a.counters = c.counters;
    return search_dispatch<true>(c, a);
}

}  // namespace sphgpu
