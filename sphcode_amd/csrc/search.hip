// Neighbour search (BHTree::neighbor_search, src/bhtree.cpp:114-126,
// 234-270) for 64 Morton-consecutive particles per wave, in three phases:
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
//  2. Candidate staging.  The collected leaves' particles (contiguous
//     ranges in key order) are read 64 at a time with coalesced loads; those
//     within the group's largest radius of the group box are kept, in
//     staging order, in the wave's staging buffer.
//  3. Candidate tiles.  With nc kept candidates and T = ceil(nc / 64)
//     tiles, tile t takes candidates t, t + T, t + 2T, ... (lane k of the
//     tile holds candidate t + kT).  Every lane applies the reference's
//     leaf test to each of the tile's candidates: |x_i - x_j|^2 < H_i^2
//     with H_i = h_guess * kernel_ratio (gather, src/pre_interaction.cpp:
//     61-70) or max(h_i, leaf's largest h) (symmetric, src/bhtree.cpp:237,
//     256), into a 64-bit hit mask, written with the tile's candidate
//     indices to the group's slab (layout: Ctx::tile_cand).
//
// The sets are the reference's; the reference then sorts each list by
// distance (src/bhtree.cpp:120-124) and stops at the first r >= h, whereas
// the pair passes here test r < h on the unsorted set -- the same set.
//
// The combined search (split mode) walks with sml = the gather radius and
// the cells' largest gather radius, keeps pairs with r < max(sml_i, sml_j),
// and marks those with r^2 < sml_i^2 (the gather candidates) in the front
// mask and the rest in the back mask, so that one walk serves both the
// pre-interaction and the force stage.
#include "ctx.hpp"

#include <type_traits>

namespace sphgpu {

struct SearchArgs {
    int *cand;           // group slabs of candidate tiles (Ctx::tile_cand)
    uint64_t *mf, *ma;   // front / all hit masks per (tile, lane)
    int *gnc;            // candidates per group
    int wcap;            // slots per group slab and per staging buffer
    int *ws_j;           // staging buffers, one per persistent wave
    double4 *ws_p;       // the staged candidate's position and (IJ) radius squared
    int split;           // symmetric search that also marks the gather set:
                         // r < sml_i in the front mask, the rest in the back
    int cap;             // per-particle capacity (the reference's list size)
    int guess;           // 1: set sml = h_guess * ratio first
    double ratio;
    double nngb;
    int64_t *counters;
    int stats;            // add to the statistics counters (errors are always counted)
    int *work;           // per-XCD task counters, zeroed before the launch
};

constexpr int SW = BLOCK / WAVE;  // waves per block
constexpr int LEAFCAP = 128;      // leaves collected per flush

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
    // the tile's candidates: position and (IJ) radius squared
    __shared__ double4 c_p[SW][WAVE];

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
            h = sml_guess<D>(m, rho, a.nngb, a.ratio);
            s.f[F_SML][i] = h;
        } else {
            h = s.f[F_SML][i];
        }
    }
    const double h2_own = h * h;
    const int64_t h2_bits = __double_as_longlong(h2_own);
    // this wave's staging buffer, and the group's slab of tiles
    const size_t wbase = (size_t)(blockIdx.x * SW + w) * (size_t)a.wcap;
    const size_t gbase = (size_t)grp * (size_t)a.wcap;

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

    int count = 0, count2 = 0;  // hits (front / back)
    int nkept = 0;               // staged candidates (may pass wcap: overflow)

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
            visited += __popcll(ballot(vis));
            const bool hit = vis && leafhit;
            const uint64_t hits = ballot(hit);
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

        // ---- phase 2: stage the collected leaves' particles, 64 at a time.
        // The next tile's candidates are located and their positions loaded
        // before the current tile is tested, so the loads overlap the work.
        int lc = 0;  // last leaf starting at or before the located tile
        auto locate = [&](int tb, int &lo_, int &j_, double xj_[D], double &lim_, bool &has_) {
            // each lane's candidate f = tb + lane; its leaf is the last leaf
            // starting at or before f, found among the few leaves that start
            // inside the tile
            const int f = tb + lane;
            int lo = lc, nx = lc;
            for(int L = lc + 1; L < nleaf; ++L) {
                const int st = uniform(l_pre[w][L]);
                if(st > tb + WAVE) break;
                nx = L;
                lo = f >= st ? L : lo;
            }
            lc = nx;
            has_ = f < ncand;
            lo_ = lo;
            j_ = has_ ? l_begin[w][lo] + (f - l_pre[w][lo]) : 0;
            if(has_) {
#pragma unroll
                for(int k = 0; k < D; ++k) xj_[k] = s.f[F_X + k][j_];
                // split mode: r < max(sml_i, sml_j) (each radius bounds that
                // particle's final h unless truncated); otherwise the
                // reference's r < max(h_i, leaf's largest h).  As squares:
                // max(a, b)^2 == max(a^2, b^2) for a, b >= 0.
                if constexpr(IJ) lim_ = a.split ? s.f[F_SML][j_] : l_ks[w][lo];
            }
        };
        int c_lo = 0, c_jj = 0;
        double c_xj[D], c_lm = 0.0;
        bool c_has = false;
        if(ncand > 0) locate(0, c_lo, c_jj, c_xj, c_lm, c_has);
        for(int tb = 0; tb < ncand; tb += WAVE) {
            int n_lo = 0, n_jj = 0;
            double n_xj[D], n_lm = 0.0;
            bool n_has = false;
            if(tb + WAVE < ncand) locate(tb + WAVE, n_lo, n_jj, n_xj, n_lm, n_has);
            // keep a candidate only if it lies within the group's largest
            // radius of the group box (the box contains every x_i, so each
            // lane's own test implies this)
            bool keep = false;
            if(c_has) {
                double d2b = 0.0;
#pragma unroll
                for(int k = 0; k < D; ++k) {
                    const double ex = fabs(min_image1(box, k, c_xj[k] - gc[k])) - ge[k];
                    d2b += ex > 0.0 ? ex * ex : 0.0;
                }
                double Hc = hmax;
                if constexpr(IJ) {
                    const double ksj = l_ks[w][c_lo];
                    Hc = ksj > Hc ? ksj : Hc;
                }
                keep = d2b <= Hc * Hc * (1.0 + 1e-9);
            }
            const uint64_t kept = ballot(keep);
            if(keep) {
                const int at = nkept + __popcll(kept & ((1ull << lane) - 1ull));
                if(at < a.wcap) {
                    a.ws_j[wbase + at] = c_jj;
                    // the position and radius are staged only where phase 3
                    // cannot read them back from the particle (the leaf's
                    // largest h of an unsplit symmetric search)
                    if(IJ && !a.split) a.ws_p[wbase + at] = make_double4(c_xj[0], D > 1 ? c_xj[1] : 0.0, D > 2 ? c_xj[2] : 0.0, c_lm * c_lm);
                }
            }
            const int m = uniform(__popcll(kept));
            nkept += m;
            tested += m;
            c_lo = n_lo;
            c_jj = n_jj;
#pragma unroll
            for(int k = 0; k < D; ++k) c_xj[k] = n_xj[k];
            c_lm = n_lm;
            c_has = n_has;
        }
        wave_lds_sync();  // leaf tables are rewritten by the next batch
    }

    // ---- phase 3: strided tiles of the kept candidates, the next tile's
    // records loaded while the current one is tested
    const int nc = nkept < a.wcap ? nkept : a.wcap;
    const int T = (nc + WAVE - 1) / WAVE;
    // the staging writes of all lanes before their reads by other lanes
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    int p_j = -1;
    double4 p_p = make_double4(0.0, 0.0, 0.0, 0.0);
    // a kept candidate's position and radius squared: staged, or read back
    // from the particle arrays (mostly still in L2 from phase 2)
    const bool staged_p = IJ && !a.split;
    auto fetch_p = [&](int j) -> double4 {
        double4 q;
        q.x = s.f[F_X][j];
        q.y = D > 1 ? s.f[F_X + 1][j] : 0.0;
        q.z = D > 2 ? s.f[F_X + 2][j] : 0.0;
        if constexpr(IJ) {
            const double l = s.f[F_SML][j];
            q.w = l * l;
        } else {
            q.w = 0.0;
        }
        return q;
    };
    if(T > 0 && lane * T < nc) {
        p_j = a.ws_j[wbase + lane * T];
        p_p = staged_p ? a.ws_p[wbase + lane * T] : fetch_p(p_j);
    }
    for(int tt = 0; tt < T; ++tt) {
        const int m = (nc - tt + T - 1) / T;  // candidates of this tile (a prefix of its lanes)
        const int j = p_j;
        if(j >= 0) c_p[w][lane] = p_p;
        stream_st(a.cand + gbase + (size_t)tt * WAVE + lane, j);
        // next tile's records
        const int cn = tt + 1 + lane * T;
        p_j = -1;
        if(tt + 1 < T && cn < nc) {
            p_j = a.ws_j[wbase + cn];
            p_p = staged_p ? a.ws_p[wbase + cn] : fetch_p(p_j);
        }
        wave_lds_sync();
        // every lane tests every candidate of the tile (uniform loop,
        // broadcast LDS reads) into a hit mask, and a mask of the hits
        // inside its own radius (the gather entries of a split search;
        // every hit otherwise).  Two candidates per pass, their records
        // read together (one with periodic boundaries); the slots past the
        // tile's m candidates hold stale records, whose bits are cleared.
        // (Four per pass spilled registers at this kernel's budget and was
        // no faster.)
        uint64_t mine = 0, mfront = 0;
        if(valid) {
            auto sweep = [&](auto periodic) {
#pragma clang fp contract(off)
                constexpr bool PER = decltype(periodic)::value;
                constexpr int U = PER ? 1 : 2;
                uint64_t hm = 0, fm = 0;
                for(int cc = 0; cc < m; cc += U) {
                    double4 q[U];
#pragma unroll
                    for(int u = 0; u < U; ++u) q[u] = c_p[w][cc + u];
                    uint32_t hn = 0, fn = 0;
#pragma unroll
                    for(int u = 0; u < U; ++u) {
                        double d[D];
                        d[0] = xi[0] - q[u].x;
                        if constexpr(D > 1) d[1] = xi[1] - q[u].y;
                        if constexpr(D > 2) d[2] = xi[2] - q[u].z;
                        if constexpr(PER) min_image<D>(box, d);
                        double r2 = d[0] * d[0];
#pragma unroll
                        for(int k = 1; k < D; ++k) r2 = r2 + d[k] * d[k];
                        // r2 < max(h_i^2, kl^2) as two tests; all three are
                        // non-negative, so their bit patterns order as integers
                        const int64_t rb = __double_as_longlong(r2);
                        const bool f = rb < h2_bits;
                        bool hh = f;
                        if constexpr(IJ) hh = f | (rb < __double_as_longlong(q[u].w));
                        hn |= (uint32_t)hh << u;
                        fn |= (uint32_t)f << u;
                    }
                    hm |= (uint64_t)hn << cc;
                    fm |= (uint64_t)fn << cc;
                }
                const uint64_t live = m >= WAVE ? ~0ull : (1ull << m) - 1ull;
                mine = hm & live;
                mfront = a.split ? fm & live : mine;
            };
            if(box.on) sweep(std::true_type{});
            else sweep(std::false_type{});
        }
        const size_t mo = gbase + (size_t)tt * WAVE + lane;
        stream_st(a.mf + mo, mfront);
        stream_st(a.ma + mo, mine);
        count += __popcll(mfront);
        count2 += __popcll(mine & ~mfront);
        wave_lds_sync();
    }
    if(lane == 0) a.gnc[grp] = nc;

    int over = 0, len = 0;
    if(valid) {
        // past the reference's list capacity (its unchecked write,
        // src/bhtree.cpp:257-258) the search reports an overflow
        over = count + count2 > a.cap ? 1 : 0;
        len = count + count2;
    }
    over = wave_sum(over);
    len = wave_maxv(len);
    if(lane == 0) {
        if(a.stats) {
            stat_add(a.counters, grp, C_SEARCH_NODES, visited);
            stat_add(a.counters, grp, C_SEARCH_CAND, tested);
            stat_max(a.counters, grp, C_MAXLIST, len);
        }
        if(over) atomicAdd((unsigned long long *)&a.counters[C_OVERFLOW], (unsigned long long)over);
        if(nkept > a.wcap) atomicAdd((unsigned long long *)&a.counters[C_TILE_OVER], 1ull);
    }
    }  // groups
}

template <int D, bool IJ> static void launch_search(Ctx &c, const SearchArgs &a, bool work_ready, int blocks_per_cu)
{
    const BoxDev box = make_box(c);
    // blocks_per_cu blocks of 4 waves per CU (8 fill every CU), a multiple
    // of the XCD count; blocks beyond what the CUs hold find no task and leave
    const int blocks = std::max(NUM_XCD, std::min(grid_for(c.n, WAVE) / SW, blocks_per_cu * c.num_cu) / NUM_XCD * NUM_XCD);
    if(!work_ready) reset_work(c, 0);
    SearchArgs aw = a;
    aw.work = c.work;
    hipLaunchKernelGGL((search_kernel<D, IJ>), dim3(blocks), dim3(BLOCK), 0, c.stream, c.s, c.t, box, c.groups, c.num_groups, aw);
}

template <bool IJ> static int search_dispatch(Ctx &c, const SearchArgs &a, bool work_ready = false, int blocks_per_cu = 8)
{
    switch(c.D) {
    case 1: launch_search<1, IJ>(c, a, work_ready, blocks_per_cu); break;
    case 2: launch_search<2, IJ>(c, a, work_ready, blocks_per_cu); break;
    default: launch_search<3, IJ>(c, a, work_ready, blocks_per_cu); break;
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
    s.f[F_SML][i] = sml_guess<D>(s.f[F_MASS][i], s.f[F_DENS][i], nngb, ratio);
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

static SearchArgs base_args(Ctx &c)
{
    SearchArgs a{};
    a.cand = c.tile_cand;
    a.mf = c.tile_mf;
    a.ma = c.tile_ma;
    a.gnc = c.tile_nc;
    a.wcap = c.tile_wcap;
    a.ws_j = c.ws_j;
    a.ws_p = c.ws_p;
    a.cap = c.cap;
    a.counters = c.counters;
    a.stats = c.stats_on ? 1 : 0;
    return a;
}

// One walk for both sets: with sml = h_guess * ratio and the cells' largest
// sml from tree_set_kernel, the test r < max(sml_i, sml_j) finds every
// gather candidate (r < sml_i, in the front masks) and every pair with
// r < max(h_i, h_j) for the final h, as long as no final h exceeds its
// search radius (the caller checks the truncation counter).  With
// `work_ready` the caller has already cleared the task counters
// (reset_work(c, 0)); `blocks_per_cu` sets how many of the CUs' wave slots
// the search's persistent waves take (8: all of them).
int search_combined(Ctx &c, bool work_ready, int blocks_per_cu)
{
    SearchArgs a = base_args(c);
    a.split = 1;
    a.guess = 0;
    return search_dispatch<true>(c, a, work_ready, blocks_per_cu);
}

int search_gather(Ctx &c, bool initial)
{
    SearchArgs a = base_args(c);
    a.guess = 1;
    a.ratio = initial ? 1.0 : (c.p.iterative_sml ? 1.2 : 1.0);  // src/pre_interaction.cpp:31-35
    a.nngb = (double)c.p.neighbor_number;
    return search_dispatch<false>(c, a, false, c.D < 3 ? 2 : 8);  // blocks per CU: see stage_pre
}

int search_symmetric(Ctx &c)
{
    SearchArgs a = base_args(c);
    a.guess = 0;
    return search_dispatch<true>(c, a, false, c.D < 3 ? 2 : 8);
}

// Slabs of candidate tiles for the c.ngroups groups of the last tree, and
// the staging buffers of the search's persistent waves.  A slab holds
// tile_wcap candidates: the larger of eight times the per-particle capacity
// and 48 times neighborNumber.  A group of 64 particles keeps about 19
// times neighborNumber in 3D (1,220 at 64 neighbours in the Evrard
// benchmark).
int search_alloc(Ctx &c)
{
    const int64_t want = std::max<int64_t>({(int64_t)8 * c.cap, (int64_t)48 * c.p.neighbor_number, 1024});
    const int wcap = (int)std::min<int64_t>((want + WAVE - 1) / WAVE * WAVE, 1 << 22);
    const int64_t waves = (int64_t)8 * c.num_cu * SW + SW;
    if(wcap != c.tile_wcap || waves > c.ws_waves) {
        if(c.ws_j) (void)hipFree(c.ws_j);
        if(c.ws_p) (void)hipFree(c.ws_p);
        c.ws_j = nullptr;
        c.ws_p = nullptr;
        SPHGPU_HIP(hipMalloc(&c.ws_j, (size_t)waves * wcap * sizeof(int)));
        SPHGPU_HIP(hipMalloc(&c.ws_p, (size_t)waves * wcap * sizeof(double4)));
        c.ws_waves = waves;
    }
    if(wcap != c.tile_wcap || c.ngroups > c.tile_groups_cap) {
        // grow with headroom: the group count changes a little every step
        const int64_t groups = std::max<int64_t>((int64_t)c.ngroups + c.ngroups / 4 + 64, c.tile_groups_cap);
        if(c.tile_cand) (void)hipFree(c.tile_cand);
        if(c.tile_mf) (void)hipFree(c.tile_mf);
        if(c.tile_ma) (void)hipFree(c.tile_ma);
        if(c.tile_nc) (void)hipFree(c.tile_nc);
        c.tile_cand = nullptr;
        c.tile_mf = c.tile_ma = nullptr;
        c.tile_nc = nullptr;
        const size_t slots = (size_t)groups * wcap;
        SPHGPU_HIP(hipMalloc(&c.tile_cand, slots * sizeof(int)));
        SPHGPU_HIP(hipMalloc(&c.tile_mf, slots * sizeof(uint64_t)));
        SPHGPU_HIP(hipMalloc(&c.tile_ma, slots * sizeof(uint64_t)));
        SPHGPU_HIP(hipMalloc(&c.tile_nc, groups * sizeof(int)));
        c.tile_groups_cap = groups;
    }
    c.tile_wcap = wcap;
    return SPHGPU_OK;
}

// Lists in list_at layout from the tiles of the last search: lane l of a
// group appends its front hits from slot 0 on and its back hits from slot
// cap - 1 down, both in staging order (candidate c = t + kT is bit k of
// tile t), which is the order the reference's one-leaf-at-a-time walk
// finds them in.  cnt / cnt2 receive the counts (cnt2 may be null with
// front_only).  Entries past cap are dropped (the search reports them).
__global__ void __launch_bounds__(BLOCK) expand_kernel(const int2 *groups, int ngroups, const int *gnc, const int *cand,
                                                       const uint64_t *mf, const uint64_t *ma, int wcap, int cap,
                                                       int front_only, int *list, int *cnt, int *cnt2)
{
    const int grp = xcd_block() * (BLOCK / WAVE) + (int)(threadIdx.x / WAVE);
    if(grp >= ngroups) return;
    const int lane = wave_lane();
    const int2 gr = groups[grp];
    const int i = gr.x + lane;
    if(i >= gr.y) return;
    const size_t gbase = (size_t)grp * wcap;
    const int nc = gnc[grp];
    const int T = (nc + WAVE - 1) / WAVE;
    const size_t lb = list_base(i, cap);
    int wf = 0, wb = 0;  // entries written (at most cap in total)
    for(int k = 0; k < WAVE; ++k) {
        for(int t = 0; t < T; ++t) {
            if(t + k * T >= nc) break;
            const size_t o = gbase + (size_t)t * WAVE;
            const uint64_t bit = 1ull << k;
            const bool room = wf + wb < cap;
            if(mf[o + lane] & bit) {
                if(room) list[lb + list_off(wf++)] = cand[o + k];
            } else if(!front_only && (ma[o + lane] & bit)) {
                if(room) list[lb + list_off(cap - 1 - wb++)] = cand[o + k];
            }
        }
    }
    cnt[i] = wf;
    if(cnt2) cnt2[i] = wb;
}

int search_expand(Ctx &c, int *list, int *cnt, int *cnt2, bool front_only)
{
    hipLaunchKernelGGL(expand_kernel, dim3(group_grid(c)), dim3(BLOCK), 0, c.stream, c.groups, c.ngroups, c.tile_nc,
                       c.tile_cand, c.tile_mf, c.tile_ma, c.tile_wcap, c.cap, front_only ? 1 : 0, list, cnt, cnt2);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

}  // namespace sphgpu
