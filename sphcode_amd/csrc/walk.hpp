// Wavefront-coherent, stackless octree walk.
//
// Nodes are stored in depth-first pre-order with a skip index (first node
// after the subtree), so one particle's traversal is a strictly increasing
// sequence of node indices: node+1 when the node is opened, skip otherwise.
// A wave of 64 particles walks together: at every iteration it takes the
// smallest cursor among its lanes, loads that node once (uniform, scalar
// loads) and lets exactly the lanes whose cursor points there apply their
// own test.  Each lane therefore makes the same per-particle decisions as
// the reference's recursive walks (src/bhtree.cpp:234-270, 301-331) while
// the wave shares node fetches and leaf particle fetches.  Every lane's
// cursor strictly increases, so the loop ends after at most num_nodes
// iterations.
#pragma once

#include "ctx.hpp"

namespace sphgpu {

constexpr int WALK_DONE = 0x7fffffff;

// Visitor interface:
//   bool open(int node)                  lane test (active lanes only)
//   void far(int node)                   node accepted without opening
//   void leaf(int4 rng, bool open)       whole wave, uniform particle loop
template <class V> __device__ __forceinline__ void tree_walk(const TreeDev &t, int total, bool valid, V &v)
{
    int cur = valid ? 0 : WALK_DONE;
    int node = uniform(wave_min_int(cur));
    while(node < total) {
        const int4 r = t.rng[node];
        const bool act = cur == node;
        const bool leaf = (r.w & 256) != 0;
        const int skip = r.z > node ? r.z : node + 1;  // cursors only move forward
        bool open = false;
        if(act) {
            open = v.open(node);
            if(!open) v.far(node);
        }
        if(leaf) {
            if(wave_any(open)) v.leaf(r, open);
            if(act) cur = skip;
        } else if(act) {
            cur = open ? node + 1 : skip;
        }
        if(wave_any(cur == node + 1)) node = node + 1;
        else node = uniform(wave_min_int(cur));
    }
}

}  // namespace sphgpu
