// Wavefront-coherent, stackless-order octree walk with per-lane decisions.
//
// Nodes are stored in depth-first pre-order with a skip index (first node
// after the subtree), so one particle's traversal is a strictly increasing
// sequence of node indices: node+1 when a node is opened, skip otherwise.
// A wave of 64 particles walks together with a uniform cursor and a mask
// of the lanes whose own traversal is at that node.  At each node those
// lanes apply their own test (as the reference's recursive walk does,
// src/bhtree.cpp:301-331); if some open it and some do not, the wave
// descends with the opening lanes and parks the others, with their mask,
// at the node's skip index on a small uniform stack.  When the descending
// lanes leave the subtree they arrive at exactly that index and the parked
// lanes rejoin.  Parked entries are sorted (the top holds the smallest
// index) and there is at most one per tree level, so the stack depth is
// bounded by the tree depth.  Every lane's cursor strictly increases, so
// the walk ends after at most num_nodes iterations per stack entry.
#pragma once

#include "ctx.hpp"

namespace sphgpu {

constexpr int WALK_DONE = 0x7fffffff;
constexpr int WALK_STACK = 72;   // > deepest possible tree (64 levels)

// Visitor interface:
//   bool open(int node)                  lane test (active lanes only)
//   void far(int node)                   node accepted without opening
//   void leaf(int4 rng, bool open)       whole wave, uniform particle loop
// Returns the number of node visits of the wave (wave-uniform).
template <class V>
__device__ __forceinline__ int tree_walk(const TreeDev &t, int total, bool valid, V &v, int *stk_node, uint64_t *stk_mask)
{
    int visits = 0;
    const int lane = wave_lane();
    uint64_t active = __ballot(valid);
    int node = 0;
    int sp = 0;                    // entries below the top, in LDS
    int top_node = WALK_DONE;      // top entry, kept in scalar registers
    uint64_t top_mask = 0;
    while(true) {
        if(active == 0 || node >= total) {
            if(top_node == WALK_DONE) break;
            node = top_node;
            active = top_mask;
            if(sp > 0) {
                --sp;
                top_node = stk_node[sp];
                top_mask = stk_mask[sp];
            } else {
                top_node = WALK_DONE;
                top_mask = 0;
            }
            continue;
        }
        if(top_node == node) {  // parked lanes rejoin
            active |= top_mask;
            if(sp > 0) {
                --sp;
                top_node = stk_node[sp];
                top_mask = stk_mask[sp];
            } else {
                top_node = WALK_DONE;
                top_mask = 0;
            }
            continue;
        }
        node = uniform(node);
        ++visits;
        const int4 r = t.rng[node];
        const bool leaf = (r.w & 256) != 0;
        const int skip = r.z > node ? r.z : node + 1;
        const bool me = (active >> lane) & 1ull;
        bool open = false;
        if(me) {
            open = v.open(node);
            if(!open) v.far(node);
        }
        const uint64_t om = __ballot(open);
        if(leaf) {
            if(om) v.leaf(r, open);
            node = skip;
        } else if(om) {
            const uint64_t fm = active & ~om;
            if(fm) {
                // park the lanes that accepted this node at its skip index
                if(top_node == skip) {
                    top_mask |= fm;
                } else {
                    if(top_node != WALK_DONE && sp < WALK_STACK) {
                        stk_node[sp] = top_node;
                        stk_mask[sp] = top_mask;
                        ++sp;
                    }
                    top_node = skip;
                    top_mask = fm;
                }
            }
            active = om;
            node = node + 1;
        } else {
            node = skip;
        }
    }
    return visits;
}

}  // namespace sphgpu
