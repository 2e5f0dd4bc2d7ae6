// Sweep of a search group's candidate tiles by its wave, one lane per
// particle (the pair passes of the pre-interaction and the force stage).
//
// stage(t) loads tile t's neighbour records into the wave's LDS slot (every
// lane loads the record of the candidate in its own position) and returns
// this lane's hit mask of the tile; refine(m), optional, narrows the mask
// to the candidates that pass a cheap test (a cut-off the pair applies
// first anyway), each lane testing its own bits once the tile is visible
// to the whole wave; pair(k) handles candidate k of the tile for this lane.
// Each lane then walks the set bits of its mask, as a divergent loop, and
// the wave moves to the next tile when every lane is done.
//
// Refining pays where the search radius holds many more candidates than
// the pair's cut-off keeps: in 3D a candidate within 1.2 h_guess passes
// r < h about 60 % of the time, so the pair loop (which runs for as long
// as any lane has bits left) would otherwise spend 40 % of its passes on a
// distance test alone; in 2D (70 %) the extra loop costs more than it saves.
// A variant that kept two tiles resident, so that lanes waited at a tile
// boundary only when one ran a whole tile ahead, needed 8-15 more registers
// and twice the LDS; it was as fast on the 3D Evrard benchmark and 7-10 %
// slower where the pair kernels run alone (2D KHI, 3D Sedov).
#pragma once

#include "ctx.hpp"

namespace sphgpu {

struct NoRefine {
    __device__ __forceinline__ uint64_t operator()(uint64_t m) const { return m; }
};

template <class Stage, class Refine, class Pair>
__device__ __forceinline__ void tile_sweep(int ntile, Stage &&stage, Refine &&refine, Pair &&pair)
{
    for(int t = 0; t < ntile; ++t) {
        uint64_t m = stage(t);
        wave_lds_sync();
        m = refine(m);
        while(m) {
            const int k = (int)__builtin_ctzll(m);
            m &= m - 1ull;
            pair(k);
        }
        wave_lds_sync();  // every read of the slot is done before it is restaged
    }
}

template <class Stage, class Pair>
__device__ __forceinline__ void tile_sweep(int ntile, Stage &&stage, Pair &&pair)
{
    tile_sweep(ntile, stage, NoRefine{}, pair);
}

// The bits of m whose candidate passes keep(k), tested by this lane.
template <class Keep>
__device__ __forceinline__ uint64_t refine_bits(uint64_t m, Keep &&keep)
{
    uint64_t out = 0;
    while(m) {
        const int k = (int)__builtin_ctzll(m);
        m &= m - 1ull;
        if(keep(k)) out |= 1ull << k;
    }
    return out;
}

}  // namespace sphgpu
