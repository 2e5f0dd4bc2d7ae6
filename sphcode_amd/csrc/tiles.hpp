// Sweep of a search group's candidate tiles by its wave, one lane per
// particle (the pair passes of the pre-interaction and the force stage).
//
// Two tiles are resident in LDS at a time.  stage(t, slot) loads tile t's
// neighbour records into LDS slot `slot` (every lane loads the record of
// the candidate in its own position) and returns this lane's hit mask of
// the tile; pair(slot, k) handles candidate k of the tile in `slot` for
// this lane.  Each lane walks the set bits of its own masks, tile after
// tile, at its own pace; the next tile is staged into the slot of the older
// resident tile as soon as no lane has bits left in it.  So lanes are not
// held at every tile boundary by the lane with the most hits in that tile,
// only when one runs more than a tile ahead of another.
#pragma once

#include "ctx.hpp"

namespace sphgpu {

// refine(slot, m), optional, narrows this lane's mask m of the tile just
// staged in `slot` to the candidates that pass a cheap test (a cut-off the
// pair would apply first anyway).  It runs once the tile is visible to the
// whole wave, as a loop over this lane's own bits, so that the sweep's pair
// loop (which every lane runs for as long as one lane has bits left) goes
// over candidates that all do the full pair work.
struct NoRefine {
    __device__ __forceinline__ uint64_t operator()(int, uint64_t m) const { return m; }
};

template <class Stage, class Refine, class Pair>
__device__ __forceinline__ void tile_sweep(int ntile, Stage &&stage, Refine &&refine, Pair &&pair)
{
    uint64_t m0 = 0, m1 = 0;  // this lane's masks of the tiles in slots 0 and 1
    int staged = 0;           // tiles staged so far (tile t sits in slot t & 1)
    if(ntile > 0) {
        m0 = stage(0, 0);
        staged = 1;
    }
    if(ntile > 1) {
        m1 = stage(1, 1);
        staged = 2;
    }
    wave_lds_sync();
    if(ntile > 0) m0 = refine(0, m0);
    if(ntile > 1) m1 = refine(1, m1);
    int myt = 0;       // the tile this lane is in
    uint64_t mm = m0;  // its bits left there
    while(true) {
        // a lane done with its tile moves to the next resident one
        if(mm == 0 && myt + 1 < staged) {
            ++myt;
            mm = (myt & 1) ? m1 : m0;
        }
        // the older resident tile (staged - 2) is free once no lane is in
        // it: a lane there without bits has moved on above
        if(staged < ntile && ballot(myt == staged - 2) == 0) {
            wave_lds_sync();  // every read of the old slot is done
            const int slot = staged & 1;
            uint64_t m = stage(staged, slot);
            ++staged;
            wave_lds_sync();
            m = refine(slot, m);
            if(slot) m1 = m;
            else m0 = m;
            continue;
        }
        const bool live = mm != 0;
        if(ballot(live) == 0) break;  // every tile staged and swept
        if(live) {
            const int k = (int)__builtin_ctzll(mm);
            mm &= mm - 1ull;
            pair(myt & 1, k);
        }
    }
    wave_lds_sync();  // the slots are reused by the caller's next sweep
}

template <class Stage, class Pair>
__device__ __forceinline__ void tile_sweep(int ntile, Stage &&stage, Pair &&pair)
{
    tile_sweep(ntile, stage, NoRefine{}, pair);
}

// The bits of m whose candidate passes keep(slot, k), tested by this lane.
template <class Keep>
__device__ __forceinline__ uint64_t refine_bits(int slot, uint64_t m, Keep &&keep)
{
    uint64_t out = 0;
    while(m) {
        const int k = (int)__builtin_ctzll(m);
        m &= m - 1ull;
        if(keep(slot, k)) out |= 1ull << k;
    }
    return out;
}

}  // namespace sphgpu
