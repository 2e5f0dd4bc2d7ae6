// Octree build on the device (replaces BHTree::make, src/bhtree.cpp:55-107,
// and BHTree::set_kernel, src/bhtree.cpp:109-112, 206-232).
//
// The reference builds a pointer octree top-down, serially: the root cube
// is the bounding box (centre, largest extent) or the periodic box
// (src/bhtree.cpp:19-31, 59-94); a particle goes to child slot
// sum_k [x_k > c_k] << k (src/bhtree.cpp:168-175); a child's centre is the
// parent's +-edge/4 and its edge half the parent's (src/bhtree.cpp:188-196);
// a child at level L is split when it holds more than leafParticleNumber
// particles and L-1 < maxTreeLevel (src/bhtree.cpp:154-158; root level 1).
//
// Here every particle replays that descent once, with the same FP64
// comparisons and centre updates, to get a key of maxTreeLevel D-bit slot
// groups (most significant = first split).  Sorting the keys puts each
// cell's particles in one contiguous range and the cells in depth-first
// pre-order, with children in slot order -- the reference's visiting order.
// The cells that exist, and which of them are leaves, follow from counts of
// equal key prefixes; a node starting at particle b exists at level L iff
// b begins a level-L cell whose parent holds more than leafParticleNumber
// particles.  Each particle emits the nodes that start at it, so an
// exclusive scan of those counts is the pre-order numbering.  Mass moments
// and the largest h per subtree are reduced bottom-up, one tree level per
// launch (the upper levels, which hold few nodes, in one single-workgroup
// launch).  The root keeps zero mass and centre of mass,
// as the reference never accumulates them (include/bhtree.hpp:46-53).
#include <cstring>

#include "ctx.hpp"

#include <rocprim/rocprim.hpp>

namespace sphgpu {

// ------------------------------------------------------------ bounding box
template <int D> __global__ void __launch_bounds__(BLOCK) bbox_partial(State s, int n, double *partial)
{
    double lo[D], hi[D];
#pragma unroll
    for(int k = 0; k < D; ++k) { lo[k] = 1.7976931348623157e308; hi[k] = -1.7976931348623157e308; }
    for(int i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK)
#pragma unroll
        for(int k = 0; k < D; ++k) {
            const double x = s.f[F_X + k][i];
            lo[k] = x < lo[k] ? x : lo[k];
            hi[k] = x > hi[k] ? x : hi[k];
        }
    __shared__ double sh[2 * D][BLOCK / WAVE];
#pragma unroll
    for(int k = 0; k < D; ++k) {
        lo[k] = wave_minv(lo[k]);
        hi[k] = wave_maxv(hi[k]);
    }
    const int w = threadIdx.x / WAVE;
    if(wave_lane() == 0)
#pragma unroll
        for(int k = 0; k < D; ++k) { sh[k][w] = lo[k]; sh[D + k][w] = hi[k]; }
    __syncthreads();
    if(threadIdx.x == 0) {
#pragma unroll
        for(int k = 0; k < D; ++k) {
            double a = sh[k][0], b = sh[D + k][0];
            for(int j = 1; j < BLOCK / WAVE; ++j) {
                a = sh[k][j] < a ? sh[k][j] : a;
                b = sh[D + k][j] > b ? sh[D + k][j] : b;
            }
            partial[blockIdx.x * 2 * D + k] = a;
            partial[blockIdx.x * 2 * D + D + k] = b;
        }
    }
}

// Root cube: centre (max+min)/2, edge = largest extent (src/bhtree.cpp:80-94).
// One wave: lanes stride over the block partials, then a wave min/max
// (exact, so the result does not depend on the order).
template <int D> __global__ void bbox_final(const double *partial, int nblocks, double *scal)
{
#pragma clang fp contract(off)
    double lo[D], hi[D];
    for(int k = 0; k < D; ++k) { lo[k] = 1.7976931348623157e308; hi[k] = -1.7976931348623157e308; }
    for(int b = threadIdx.x; b < nblocks; b += WAVE)
        for(int k = 0; k < D; ++k) {
            const double a = partial[b * 2 * D + k], c = partial[b * 2 * D + D + k];
            lo[k] = a < lo[k] ? a : lo[k];
            hi[k] = c > hi[k] ? c : hi[k];
        }
    for(int k = 0; k < D; ++k) {
        lo[k] = wave_minv(lo[k]);
        hi[k] = wave_maxv(hi[k]);
    }
    if(threadIdx.x != 0) return;
    double edge = 0.0;
    for(int k = 0; k < D; ++k) {
        scal[S_ROOT_CX + k] = (hi[k] + lo[k]) * 0.5;
        const double r = hi[k] - lo[k];
        if(edge < r) edge = r;
    }
    scal[S_ROOT_EDGE] = edge;
}

// ------------------------------------------------------------------- keys
// The reference's descent, one level per slot group (src/bhtree.cpp:163-204).
template <int D> __global__ void __launch_bounds__(BLOCK) keys_kernel(State s, int n, const double *scal, int levels, uint64_t *keys, int *idx)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
    double c[D], x[D];
#pragma unroll
    for(int k = 0; k < D; ++k) { c[k] = scal[S_ROOT_CX + k]; x[k] = s.f[F_X + k][i]; }
    double e = scal[S_ROOT_EDGE];
    uint64_t key = 0;
    for(int L = 0; L < levels; ++L) {
        uint64_t slot = 0;
#pragma unroll
        for(int k = 0; k < D; ++k) {
            const bool up = x[k] > c[k];
            slot |= (uint64_t)up << k;
            c[k] = c[k] + (up ? 1.0 : -1.0) * e * 0.25;
        }
        e = e * 0.5;
        key = (key << D) | slot;
    }
    keys[i] = key;
    idx[i] = i;
}

// --------------------------------------------------------------- reorder
struct Reorder {
    const double *src[NF];
    double *dst[NF];
    const int *isrc[2];
    int *idst[2];
    // with sml set: the gather search radius of every particle in the new
    // order, from its mass and density (the combined search's radius)
    double *sml;
    const double *mass, *dens;
    double nngb, ratio;
};

template <int D> __global__ void __launch_bounds__(BLOCK) reorder_kernel(Reorder r, const int *perm, int n, int nfields)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if(i >= n) return;
    const int j = perm[i];
    for(int f = 0; f < nfields; ++f) r.dst[f][i] = r.src[f][j];
    r.idst[0][i] = r.isrc[0][j];
    if(r.isrc[1]) r.idst[1][i] = r.isrc[1][j];
    if(r.sml) r.sml[i] = sml_guess<D>(r.mass[j], r.dens[j], r.nngb, r.ratio);
}

// ----------------------------------------------------------------- delta
// delta[i] = number of equal leading slot groups of keys i-1 and i: the two
// particles share the cells of levels 1 .. delta+1.
__device__ __forceinline__ int key_delta(const uint64_t *k, int i, int G, int D)
{
    if(i == 0) return 0;
    const uint64_t x = k[i - 1] ^ k[i];
    if(x == 0) return G;
    const int lz = __clzll((long long)x) - (64 - G * D);
    return lz / D;
}

// Levels of the per-level node lists (> deepest level: max_level + 1 <= 65).
constexpr int NLEVEL = 72;

// Search groups: compact sets of at most 64 key-consecutive particles.
// Every cell holding more than 64 particles packs its children that hold at
// most 64 (cells the tree does not split further, or leaves) into groups,
// consecutive children filling a group until the next would take it past
// 64; a child holding more than 64 is packed by its own children, and a
// leaf holding more than 64 is cut into runs of 64.  A group then lies in
// one parent cell, so its box is compact, and it is as full as siblings
// allow (46.7 particles on average at the Evrard benchmark's density
// profile, against 15.6 for single cells; runs of 64 key-consecutive
// particles fill the waves but can straddle a jump of the key order across
// the domain).  In 1D and 2D a stretch of consecutive small children is
// instead cut into equal groups of at most 64 (still inside the parent).
constexpr int GROUP = 64;

struct TreeShape {
    int n, G, D, leaf, max_level;
};

// First level at which particle b begins a cell, from its delta.
__device__ __forceinline__ int first_level(int delta_b, int b) { return b == 0 ? 2 : delta_b + 2; }

// Level-L cells are the runs of sorted keys that share their top L-1 levels
// (the prefix cell_end compares).  The two helpers below find a run's ends
// within leaf positions of b by binary search over the keys, so every lane
// takes the same few steps (a scan of the key-difference bytes took up to
// 2 leaf steps, different for each lane).
__device__ __forceinline__ uint64_t cell_prefix(uint64_t k, int shift) { return shift >= 64 ? 0 : k >> shift; }

// Size of the level-L cell that begins at b, or leaf+1 if it is larger.
__device__ __forceinline__ int cell_count_bounded(const uint64_t *keys, const TreeShape &ts, int b, int L)
{
    const int shift = (ts.G - (L - 1)) * ts.D;
    const uint64_t pre = cell_prefix(keys[b], shift);
    // last member of the run in [b, b + leaf]
    int lo = b, hi = min(ts.n - 1, b + ts.leaf);
    while(lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if(cell_prefix(keys[mid], shift) > pre) hi = mid - 1;
        else lo = mid;
    }
    return lo - b + 1;  // leaf + 1 when the run reaches b + leaf
}

// Does the level-P cell containing b hold more than leaf particles?  Its
// members within leaf positions on either side of b number more than leaf
// exactly when the whole cell does.
__device__ __forceinline__ bool parent_split(const uint64_t *keys, const TreeShape &ts, int b, int P)
{
    const int shift = (ts.G - (P - 1)) * ts.D;
    const uint64_t pre = cell_prefix(keys[b], shift);
    int lo = max(0, b - ts.leaf), hi = b;  // first member in [b - leaf, b]
    while(lo < hi) {
        const int mid = (lo + hi) >> 1;
        if(cell_prefix(keys[mid], shift) < pre) lo = mid + 1;
        else hi = mid;
    }
    const int first = lo;
    lo = b;
    hi = min(ts.n - 1, b + ts.leaf);  // last member in [b, b + leaf]
    while(lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if(cell_prefix(keys[mid], shift) > pre) hi = mid - 1;
        else lo = mid;
    }
    return lo - first + 1 > ts.leaf;
}

// Number of nodes that begin at particle b (all at consecutive levels).
__device__ int nodes_at(const uint64_t *keys, int delta_b, const TreeShape &ts, int b, int &lvl0)
{
    if(b > 0 && delta_b >= ts.G) return 0;  // same deepest cell as b-1
    lvl0 = first_level(delta_b, b);
    if(lvl0 > 2 && !parent_split(keys, ts, b, lvl0 - 1)) return 0;
    int k = 0;
    for(int L = lvl0;; ++L) {
        ++k;
        const int cnt = cell_count_bounded(keys, ts, b, L);
        if(!(cnt > ts.leaf && L <= ts.max_level)) break;
    }
    return k;
}

// Nodes beginning at each particle, and the deltas for emit_nodes_kernel.
__global__ void __launch_bounds__(BLOCK) count_nodes_kernel(const uint64_t *keys, TreeShape ts, uint8_t *delta, int *nn)
{
    const int b = blockIdx.x * BLOCK + threadIdx.x;
    if(b >= ts.n) return;
    const int d = key_delta(keys, b, ts.G, ts.D);
    delta[b] = (uint8_t)d;
    int lvl0 = 0;
    nn[b] = nodes_at(keys, d, ts, b, lvl0);
}

// End (exclusive) of the level-L cell that begins at b: first particle
// whose key differs in the first L-1 slot groups.
__device__ __forceinline__ int cell_end(const uint64_t *keys, const TreeShape &ts, int b, int L)
{
    const int shift = (ts.G - (L - 1)) * ts.D;
    const uint64_t pre = shift >= 64 ? 0 : keys[b] >> shift;
    int lo = b + 1, hi = ts.n;
    while(lo < hi) {
        const int mid = (lo + hi) >> 1;
        const uint64_t pm = shift >= 64 ? 0 : keys[mid] >> shift;
        if(pm > pre) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Also clears the per-level counters (lvl) and the search groups: the
// groups of a root holding at most GROUP particles, none otherwise.
template <int D>
__global__ void __launch_bounds__(BLOCK) emit_nodes_kernel(const uint64_t *keys, const uint8_t *delta, TreeShape ts, const int *nn,
                                                        const int *noff, const double *scal, TreeDev t, int64_t node_cap,
                                                        int64_t *counters, int *lvl, int2 *groups, int *num_groups)
{
#pragma clang fp contract(off)
    const int b = blockIdx.x * BLOCK + threadIdx.x;
    if(b < NLEVEL) lvl[b] = 0;  // the level counts (the grid has at least BLOCK > NLEVEL threads)
    if(b >= ts.n) return;
    const int total = 1 + noff[ts.n - 1] + nn[ts.n - 1];
    if(b == 0) {
        *t.num_nodes = total <= node_cap ? total : (int)node_cap;
        if(total > node_cap) atomicAdd((unsigned long long *)&counters[C_TREE_ERR], 1ull);
        t.geo[0] = make_double4(scal[S_ROOT_CX], D > 1 ? scal[S_ROOT_CY] : 0.0, D > 2 ? scal[S_ROOT_CZ] : 0.0, scal[S_ROOT_EDGE]);
        t.rng[0] = make_int4(0, ts.n, total, 1);
        int ng = 0;
        if(ts.n <= GROUP) groups[ng++] = make_int2(0, ts.n);
        *num_groups = ng;
    }
    const int cnt = nn[b];
    if(cnt == 0) return;
    int lvl0 = 0;
    nodes_at(keys, (int)delta[b], ts, b, lvl0);
    const int base = 1 + noff[b];
    const uint64_t key = keys[b];
    const uint64_t mask = (1ull << D) - 1ull;
    double c[D];
#pragma unroll
    for(int k = 0; k < D; ++k) c[k] = scal[S_ROOT_CX + k];
    double e = scal[S_ROOT_EDGE];
    // replay the descent down to level lvl0 - 1
    int g = 0;
    for(; g < lvl0 - 2; ++g) {
        const uint64_t slot = (key >> ((ts.G - 1 - g) * D)) & mask;
#pragma unroll
        for(int k = 0; k < D; ++k) c[k] = c[k] + (((slot >> k) & 1) ? 1.0 : -1.0) * e * 0.25;
        e = e * 0.5;
    }
    for(int j = 0; j < cnt; ++j) {
        const int L = lvl0 + j;
        // descend one more level: cell of level L
        const uint64_t slot = (key >> ((ts.G - 1 - g) * D)) & mask;
#pragma unroll
        for(int k = 0; k < D; ++k) c[k] = c[k] + (((slot >> k) & 1) ? 1.0 : -1.0) * e * 0.25;
        e = e * 0.5;
        ++g;
        const int node = base + j;
        const bool leaf = j == cnt - 1;
        int end;
        if(leaf && L <= ts.max_level) end = b + cell_count_bounded(keys, ts, b, L);
        else end = cell_end(keys, ts, b, L);
        const int skip = end < ts.n ? 1 + noff[end] : total;
        if(node < node_cap) {
            t.geo[node] = make_double4(c[0], D > 1 ? c[D > 1 ? 1 : 0] : 0.0, D > 2 ? c[D > 2 ? 2 : 0] : 0.0, e);
            t.rng[node] = make_int4(b, end, skip, L | (leaf ? 256 : 0));
        }
    }
}

// ------------------------------------------------- level-by-level reductions
// Node sums are formed bottom-up one tree level per launch: a launch
// boundary orders the children's results before the parents read them, so
// no cross-XCD atomics or cache maintenance are needed.  The nodes of each
// level are listed once per tree build (level_bins / level_scatter).

__device__ __forceinline__ int node_level(const int4 &r) { return (r.w & 255) < NLEVEL ? (r.w & 255) : NLEVEL - 1; }

// Layout of lvl: counts [0, NLEVEL), offsets [NLEVEL, 2 NLEVEL), the number
// of search groups at 2 NLEVEL (so that one copy takes all three to the
// host), fill counters from LVL_FILL on.
constexpr int LVL_FILL = 2 * NLEVEL + 1;

// offsets (lvl[NLEVEL + L]) from the counts; clears the fill counters
__global__ void level_offsets_kernel(int *lvl)
{
    if(threadIdx.x != 0) return;
    int o = 0;
    for(int k = 0; k < NLEVEL; ++k) {
        lvl[NLEVEL + k] = o;
        lvl[LVL_FILL + k] = 0;
        o += lvl[k];
    }
}

__global__ void __launch_bounds__(BLOCK) level_scatter_kernel(TreeDev t, int *lvl, int *nodes)
{
    __shared__ int bins[NLEVEL], base[NLEVEL];
    for(int k = threadIdx.x; k < NLEVEL; k += BLOCK) bins[k] = 0;
    __syncthreads();
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    const int total = *t.num_nodes;
    int L = 0, rank = 0;
    if(p < total) {
        L = node_level(t.rng[p]);
        rank = atomicAdd(&bins[L], 1);
    }
    __syncthreads();
    for(int k = threadIdx.x; k < NLEVEL; k += BLOCK)
        base[k] = bins[k] ? atomicAdd(&lvl[LVL_FILL + k], bins[k]) : 0;
    __syncthreads();
    if(p < total) nodes[lvl[NLEVEL + L] + base[L] + rank] = p;
}

// Mass and centre of mass of the nodes of one level: a leaf sums its
// particles, an internal node its children (one level deeper, done by the
// previous launch), in child order.  The root's moments stay zero, as the
// reference never accumulates them (include/bhtree.hpp:46-53).
// With KS, also the node's largest smoothing length (kernel_size_node's
// result) in the same pass.
template <int D, bool KS> __device__ __forceinline__ void moments_node(const State &s, const TreeDev &t, int p, int total, double4 *msum)
{
    const int4 r = t.rng[p];
    double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
    double h = 0.0;
    if(r.w & 256) {
        double mx[3] = {0.0, 0.0, 0.0};
        for(int i = r.x; i < r.y; ++i) {
            const double mi = s.f[F_MASS][i];
            acc.w += mi;
#pragma unroll
            for(int q = 0; q < D; ++q) mx[q] += s.f[F_X + q][i] * mi;
            if constexpr(KS) {
                const double hi = s.f[F_SML][i];
                h = hi > h ? hi : h;
            }
        }
        acc.x = mx[0];
        acc.y = mx[1];
        acc.z = mx[2];
    } else {
        int guard = 0;
        for(int c = p + 1; c < r.z && c < total && guard < 8; c = t.rng[c].z, ++guard) {
            const double4 v = msum[c];
            acc.x += v.x; acc.y += v.y; acc.z += v.z; acc.w += v.w;
            if constexpr(KS) {
                const double hc = t.ks[c];
                h = hc > h ? hc : h;
            }
        }
    }
    if constexpr(KS) t.ks[p] = h;
    msum[p] = acc;
    if(p == 0) t.mom[0] = make_double4(0.0, 0.0, 0.0, 0.0);
    else t.mom[p] = make_double4(acc.x / acc.w, acc.y / acc.w, acc.z / acc.w, acc.w);
}

template <int D, bool KS>
__global__ void __launch_bounds__(BLOCK) moments_level_kernel(State s, TreeDev t, const int *nodes, int count, double4 *msum)
{
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if(k >= count) return;
    moments_node<D, KS>(s, t, nodes[k], *t.num_nodes, msum);
}

// The upper levels, which hold few nodes, in one launch of one workgroup:
// level after level bottom-up, a barrier between levels orders the
// children's results before their parents read them (one workgroup runs on
// one CU, whose waves see each other's stores once past the barrier).
constexpr int TOP_THREADS = 1024;
constexpr int TOP_MAXNODES = 512;  // a level is one of the upper ones up to this many nodes
struct TopLevels {
    int first;        // levels first .. first + nlev - 1 (entry k is level first + k)
    int nlev;
    int count[16];
    int offset[16];
};

template <int D, bool KS>
__global__ void __launch_bounds__(TOP_THREADS) moments_top_kernel(State s, TreeDev t, const int *nodes, TopLevels tl, double4 *msum)
{
    const int total = *t.num_nodes;
    for(int L = tl.nlev - 1; L >= 0; --L) {
        for(int k = threadIdx.x; k < tl.count[L]; k += TOP_THREADS) moments_node<D, KS>(s, t, nodes[tl.offset[L] + k], total, msum);
        __syncthreads();
    }
}


// Per node: the per-level node counts (block histograms, one atomic per bin
// and block), and the groups of a packing cell.
__global__ void __launch_bounds__(BLOCK) level_count_groups_kernel(TreeDev t, int *lvl, int2 *groups, int *num_groups,
                                                                int64_t group_cap, int runs)
{
    __shared__ int bins[NLEVEL];
    for(int k = threadIdx.x; k < NLEVEL; k += BLOCK) bins[k] = 0;
    __syncthreads();
    const int q = blockIdx.x * BLOCK + threadIdx.x;
    const int total = *t.num_nodes;
    int4 rq = make_int4(0, 0, 0, 0);
    if(q < total) {
        rq = t.rng[q];
        atomicAdd(&bins[node_level(rq)], 1);
    }
    __syncthreads();
    for(int k = threadIdx.x; k < NLEVEL; k += BLOCK)
        if(bins[k]) atomicAdd(&lvl[k], bins[k]);
    if(q >= total) return;
    if((rq.w & 256) || rq.y - rq.x <= GROUP) return;  // not a packing cell
    const int end = rq.z > q ? rq.z : total;
    // two passes over the children: count the groups, then write them
    int g0 = 0;
    for(int pass = 0; pass < 2; ++pass) {
        int ng = 0, cb = -1, ce = -1;
        auto emit = [&](int b, int e) {
            if(pass == 1 && g0 + ng < group_cap) groups[g0 + ng] = make_int2(b, e);
            ++ng;
        };
        // runs: a stretch [b, e) of consecutive small children is cut into
        // ceil((e - b) / GROUP) groups of equal size, which fills the waves
        // but does not keep whole cells together
        auto cut = [&](int b, int e) {
            const int k = (e - b + GROUP - 1) / GROUP;
            for(int j = 0; j < k; ++j) emit(b + (int)((int64_t)(e - b) * j / k), b + (int)((int64_t)(e - b) * (j + 1) / k));
        };
        for(int c = q + 1; c < end;) {
            const int4 rc = t.rng[c];
            const int cnt = rc.y - rc.x;
            const bool leaf = (rc.w & 256) != 0;
            if(cnt > GROUP) {
                if(cb >= 0) {
                    if(runs) cut(cb, ce);
                    else emit(cb, ce);
                }
                cb = -1;
                if(leaf)
                    for(int b = rc.x; b < rc.y; b += GROUP) emit(b, b + GROUP < rc.y ? b + GROUP : rc.y);
            } else if(cnt > 0) {
                if(!runs && cb >= 0 && rc.y - cb > GROUP) {
                    emit(cb, ce);
                    cb = -1;
                }
                if(cb < 0) cb = rc.x;
                ce = rc.y;
            }
            c = rc.z > c ? rc.z : c + 1;
        }
        if(cb >= 0) {
            if(runs) cut(cb, ce);
            else emit(cb, ce);
        }
        if(pass == 0) g0 = atomicAdd(num_groups, ng);
    }
}

// Largest smoothing length per subtree (BHNode::set_kernel,
// src/bhtree.cpp:206-232), one level per launch as the moments.
__device__ __forceinline__ void kernel_size_node(const State &s, const TreeDev &t, int p, int total)
{
    const int4 r = t.rng[p];
    double h = 0.0;
    if(r.w & 256) {
        for(int i = r.x; i < r.y; ++i) {
            const double hi = s.f[F_SML][i];
            h = hi > h ? hi : h;
        }
    } else {
        int guard = 0;
        for(int c = p + 1; c < r.z && c < total && guard < 8; c = t.rng[c].z, ++guard) {
            const double v = t.ks[c];
            h = v > h ? v : h;
        }
    }
    t.ks[p] = h;
}

__global__ void __launch_bounds__(BLOCK) kernel_size_level_kernel(State s, TreeDev t, const int *nodes, int count)
{
    const int k = blockIdx.x * BLOCK + threadIdx.x;
    if(k >= count) return;
    kernel_size_node(s, t, nodes[k], *t.num_nodes);
}

__global__ void __launch_bounds__(TOP_THREADS) kernel_size_top_kernel(State s, TreeDev t, const int *nodes, TopLevels tl)
{
    const int total = *t.num_nodes;
    for(int L = tl.nlev - 1; L >= 0; --L) {
        for(int k = threadIdx.x; k < tl.count[L]; k += TOP_THREADS) kernel_size_node(s, t, nodes[tl.offset[L] + k], total);
        __syncthreads();
    }
}

// ------------------------------------------------------------------ host
int tree_alloc(Ctx &c)
{
    const int64_t n = c.cap_n;
    c.node_cap = 5 * n + 1;  // BHTree::resize: 5 nodes per particle (+ root)
    SPHGPU_HIP(hipMalloc(&c.keys, n * sizeof(uint64_t)));
    SPHGPU_HIP(hipMalloc(&c.keys_sorted, n * sizeof(uint64_t)));
    SPHGPU_HIP(hipMalloc(&c.idx, n * sizeof(int)));
    SPHGPU_HIP(hipMalloc(&c.perm, n * sizeof(int)));
    SPHGPU_HIP(hipMalloc(&c.delta, n));
    SPHGPU_HIP(hipMalloc(&c.nn, n * sizeof(int)));
    SPHGPU_HIP(hipMalloc(&c.noff, n * sizeof(int)));
    SPHGPU_HIP(hipMalloc(&c.t.geo, c.node_cap * sizeof(double4)));
    SPHGPU_HIP(hipMalloc(&c.t.mom, c.node_cap * sizeof(double4)));
    SPHGPU_HIP(hipMalloc(&c.t.ks, c.node_cap * sizeof(double)));
    SPHGPU_HIP(hipMalloc(&c.t.rng, c.node_cap * sizeof(int4)));
    SPHGPU_HIP(hipMalloc(&c.t.num_nodes, sizeof(int)));
    SPHGPU_HIP(hipMalloc(&c.lvl_nodes, c.node_cap * sizeof(int)));
    SPHGPU_HIP(hipMalloc(&c.msum, c.node_cap * sizeof(double4)));
    if(c.p.gravity_is_valid) {
        SPHGPU_HIP(hipMalloc(&c.gnode, c.node_cap * sizeof(GravNode)));
        // one spare record: a scalar load may fetch a whole record past the last
        SPHGPU_HIP(hipMalloc(&c.gpart, (n + 1) * sizeof(GravPart<double>)));
        SPHGPU_HIP(hipMalloc(&c.gacc, (size_t)c.D * n * sizeof(double)));
        // one record slot per wave of a grid_for(n)-block launch
        const size_t waves = (size_t)grid_for(n) * (BLOCK / WAVE);
        SPHGPU_HIP(hipMalloc(&c.leafhit, waves * LEAFHIT_CAP * sizeof(LeafHit)));
        SPHGPU_HIP(hipMalloc(&c.leafcnt, waves * sizeof(int)));
    }
    if(!c.lvl) SPHGPU_HIP(hipMalloc(&c.lvl, (LVL_FILL + NLEVEL) * sizeof(int)));
    c.num_groups = c.lvl + 2 * NLEVEL;
    // per-level counts and offsets, then the number of search groups
    if(!c.lvl_host) SPHGPU_HIP(hipHostMalloc(&c.lvl_host, (2 * NLEVEL + 1) * sizeof(int), hipHostMallocDefault));
    for(int k = 0; k < 2 * NLEVEL + 1; ++k) c.lvl_host[k] = 0;
    c.group_cap = c.node_cap + n / 64 + 1;
    SPHGPU_HIP(hipMalloc(&c.groups, c.group_cap * sizeof(int2)));
    size_t bytes = 0;
    const unsigned key_bits = (unsigned)(c.p.tree_max_level * c.D);
    SPHGPU_HIP(rocprim::radix_sort_pairs(nullptr, bytes, c.keys, c.keys_sorted, c.idx, c.perm, (unsigned)n, 0, key_bits, c.stream));
    c.sort_tmp_bytes = bytes;
    SPHGPU_HIP(hipMalloc(&c.sort_tmp, bytes));
    bytes = 0;
    SPHGPU_HIP(rocprim::exclusive_scan(nullptr, bytes, c.nn, c.noff, 0, (size_t)n, rocprim::plus<int>(), c.stream));
    c.scan_tmp_bytes = bytes;
    SPHGPU_HIP(hipMalloc(&c.scan_tmp, bytes));
    return SPHGPU_OK;
}

// The upper levels that one workgroup reduces: levels 0 .. nlev-1, each with
// at most TOP_MAXNODES nodes (level counts from the last tree build).
static TopLevels top_levels(const Ctx &c)
{
    TopLevels tl{};
    int L = 0;
    while(L < NLEVEL && c.lvl_host[L] == 0) ++L;  // levels start at the root's
    tl.first = L;
    int k = 0;
    while(k < 16 && L + k < NLEVEL && c.lvl_host[L + k] > 0 && c.lvl_host[L + k] <= TOP_MAXNODES) {
        tl.count[k] = c.lvl_host[L + k];
        tl.offset[k] = c.lvl_host[NLEVEL + L + k];
        ++k;
    }
    tl.nlev = k;
    return tl;
}

// Fields that a sphgpu_integrate step writes after its tree build before
// any kernel reads them: acc and du/dt (force pass), pressure, h and 1/h
// (pre-interaction; h is first the search radius the guess writes),
// grad-h (SSPH, DISPH), the Balsara switch where the pre-interaction sets it,
// and phi (gravity walk).
static bool rewritten_in_step(const Ctx &c, int f)
{
    if(f >= F_A && f < F_A + 3) return true;
    if(f == F_DENE || f == F_PRES || f == F_SML || f == F_HINV) return true;
    if(f == F_GRADH) return c.p.sph_type != SPHGPU_GSPH;
    if(f == F_BAL) return c.p.av_use_balsara_switch && c.D != 1 && c.p.sph_type != SPHGPU_GSPH;
    if(f == F_PHI) return c.p.gravity_is_valid != 0;
    return false;
}

template <int D> static int tree_build_d(Ctx &c)
{
    const int n = (int)c.n;
    const int G = c.p.tree_max_level;
    c.radius_ready = false;
    // root geometry
    if(!c.p.periodic_is_valid) {
        const int nb = std::min(grid_for(n), 1024);
        hipLaunchKernelGGL(bbox_partial<D>, dim3(nb), dim3(BLOCK), 0, c.stream, c.s, n, c.partial);
        hipLaunchKernelGGL(bbox_final<D>, dim3(1), dim3(WAVE), 0, c.stream, c.partial, nb, c.scal);
    }
    hipLaunchKernelGGL(keys_kernel<D>, dim3(grid_for(n)), dim3(BLOCK), 0, c.stream, c.s, n, c.scal, G, c.keys, c.idx);
    size_t bytes = c.sort_tmp_bytes;
    SPHGPU_HIP(rocprim::radix_sort_pairs(c.sort_tmp, bytes, c.keys, c.keys_sorted, c.idx, c.perm, (unsigned)n, 0, (unsigned)(G * D), c.stream));
    // gather the state into key order (in sphgpu_integrate, without the
    // fields that the step rewrites before anything reads them)
    Reorder r{};
    int nf = 0;
    for(int f = 0; f < NF; ++f) {
        if(!c.s.f[f]) continue;
        if(c.reorder_light && rewritten_in_step(c, f)) continue;
        r.src[nf] = c.s.f[f];
        r.dst[nf] = c.alt.f[f];
        ++nf;
    }
    r.isrc[0] = c.s.id; r.idst[0] = c.alt.id;
    r.isrc[1] = c.reorder_light ? nullptr : c.s.nb; r.idst[1] = c.alt.nb;
    // sphgpu_integrate's combined search needs the radii h_guess * ratio and
    // the cells' largest radius: set here, by the reorder and the moments
    const bool radius = c.reorder_light;
    if(radius) {
        r.sml = c.alt.f[F_SML];
        r.mass = c.s.f[F_MASS];
        r.dens = c.s.f[F_DENS];
        r.nngb = (double)c.p.neighbor_number;
        r.ratio = c.p.iterative_sml ? 1.2 : 1.0;
    }
    hipLaunchKernelGGL(reorder_kernel<D>, dim3(grid_for(n)), dim3(BLOCK), 0, c.stream, r, c.perm, n, nf);
    // swap master buffers; FP32 mirrors stay (they are rewritten before use)
    for(int f = 0; f < NF; ++f) std::swap(c.s.f[f], c.alt.f[f]);
    std::swap(c.s.id, c.alt.id);
    std::swap(c.s.nb, c.alt.nb);

    const TreeShape ts{n, G, D, c.p.tree_leaf_particle_num, c.p.tree_max_level};
    hipLaunchKernelGGL(count_nodes_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, c.stream, c.keys_sorted, ts, c.delta, c.nn);
    bytes = c.scan_tmp_bytes;
    SPHGPU_HIP(rocprim::exclusive_scan(c.scan_tmp, bytes, c.nn, c.noff, 0, (size_t)n, rocprim::plus<int>(), c.stream));
    hipLaunchKernelGGL(emit_nodes_kernel<D>, dim3(grid_for(n)), dim3(BLOCK), 0, c.stream, c.keys_sorted, c.delta, ts, c.nn, c.noff,
                       c.scal, c.t, c.node_cap, c.counters, c.lvl, c.groups, c.num_groups);
    // nodes listed by level (in c.lvl_nodes), and the search groups (they
    // need the node ranges only), so that the level counts and the number of
    // groups come to the host together
    const int ngrid = grid_for(c.node_cap);
    // in 1D and 2D the waves are filled from runs of small cells (2D KHI
    // step 2 % faster); in 3D whole cells stay together (the Evrard step
    // was 6 % slower with runs: less compact groups keep more candidates)
    hipLaunchKernelGGL(level_count_groups_kernel, dim3(ngrid), dim3(BLOCK), 0, c.stream, c.t, c.lvl, c.groups, c.num_groups,
                       c.group_cap, D < 3 ? 1 : 0);
    hipLaunchKernelGGL(level_offsets_kernel, dim3(1), dim3(WAVE), 0, c.stream, c.lvl);
    hipLaunchKernelGGL(level_scatter_kernel, dim3(ngrid), dim3(BLOCK), 0, c.stream, c.t, c.lvl, c.lvl_nodes);
    SPHGPU_HIP(hipMemcpyAsync(c.lvl_host, c.lvl, (2 * NLEVEL + 1) * sizeof(int), hipMemcpyDeviceToHost, c.stream));
    SPHGPU_HIP(hipStreamSynchronize(c.stream));
    c.ngroups = c.lvl_host[2 * NLEVEL];
    if(c.ngroups < 1 || c.ngroups > c.group_cap) return set_error(SPHGPU_E_TREE, "search groups out of range");
    const TopLevels tl = top_levels(c);
    for(int L = NLEVEL - 1; L >= tl.first + tl.nlev; --L) {
        const int cnt = c.lvl_host[L];
        if(cnt <= 0) continue;
        if(radius)
            hipLaunchKernelGGL((moments_level_kernel<D, true>), dim3(grid_for(cnt)), dim3(BLOCK), 0, c.stream, c.s, c.t,
                               c.lvl_nodes + c.lvl_host[NLEVEL + L], cnt, c.msum);
        else
            hipLaunchKernelGGL((moments_level_kernel<D, false>), dim3(grid_for(cnt)), dim3(BLOCK), 0, c.stream, c.s, c.t,
                               c.lvl_nodes + c.lvl_host[NLEVEL + L], cnt, c.msum);
    }
    if(tl.nlev > 0) {
        if(radius)
            hipLaunchKernelGGL((moments_top_kernel<D, true>), dim3(1), dim3(TOP_THREADS), 0, c.stream, c.s, c.t, c.lvl_nodes, tl, c.msum);
        else
            hipLaunchKernelGGL((moments_top_kernel<D, false>), dim3(1), dim3(TOP_THREADS), 0, c.stream, c.s, c.t, c.lvl_nodes, tl, c.msum);
    }
    SPHGPU_HIP(hipGetLastError());
    c.radius_ready = radius;
    return search_alloc(c);
}

int tree_build(Ctx &c)
{
    switch(c.D) {
    case 1: return tree_build_d<1>(c);
    case 2: return tree_build_d<2>(c);
    default: return tree_build_d<3>(c);
    }
}

int tree_set_kernel(Ctx &c)
{
    // level lists and counts from the last tree build
    const TopLevels tl = top_levels(c);
    for(int L = NLEVEL - 1; L >= tl.first + tl.nlev; --L) {
        const int cnt = c.lvl_host[L];
        if(cnt <= 0) continue;
        hipLaunchKernelGGL(kernel_size_level_kernel, dim3(grid_for(cnt)), dim3(BLOCK), 0, c.stream, c.s, c.t,
                           c.lvl_nodes + c.lvl_host[NLEVEL + L], cnt);
    }
    if(tl.nlev > 0)
        hipLaunchKernelGGL(kernel_size_top_kernel, dim3(1), dim3(TOP_THREADS), 0, c.stream, c.s, c.t, c.lvl_nodes, tl);
    SPHGPU_HIP(hipGetLastError());
    return SPHGPU_OK;
}

}  // namespace sphgpu
