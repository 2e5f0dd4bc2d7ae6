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

// Visitor interface (every call is made by the whole wave):
//   uint64_t node(const GravNode &g, uint64_t act)
//                                        the opening decisions at the node of
//                                        the lanes in `act` (the lanes whose
//                                        traversal is there), as the mask of
//                                        the lanes that open it; the visitor
//                                        sums the node for the lanes of `act`
//                                        that do not open it, masked rather
//                                        than branched
//   void leaf(int b, int e, uint64_t open)
//                                        particles [b, e) of an opened leaf,
//                                        for the lanes of `open`
// The walk state (cursor, lane masks, parked stack) is wave-uniform and kept
// in scalar registers; lanes take their part of a mask with lane_in, so no
// per-lane flag is carried from one visit to the next.  Each node is one
// 64-byte record (GravNode) read with one scalar load.  With STATS,
// WalkCount returns the wave's node visits, the particles of the leaves
// that some lane opened, and the lanes' opened nodes, accepted nodes and
// particle interactions, all wave totals counted from the lane masks with
// scalar bit counts (no vector instructions per visit, but five scalar
// ones: the gravity stage alone took 2.85 against 2.66 ms without them at
// the Evrard benchmark, so they are counted only on request).
struct WalkCount {
    int visits = 0, slots = 0;
    int opened = 0, accepted = 0, pairs = 0;
};

template <bool STATS, class V>
__device__ __forceinline__ WalkCount tree_walk(const GravNode *gn, int total, bool valid, V &v, int *stk_node,
                                               uint64_t *stk_mask)
{
    WalkCount wc;
    uint64_t active = ballot(valid);
    int node = 0;
    int sp = 0;                    // entries below the top, in LDS
    int top_node = WALK_DONE;      // top entry
    uint64_t top_mask = 0;
    auto pop = [&]() {
        if(sp > 0) {
            --sp;
            top_node = uniform(stk_node[sp]);
            top_mask = uniform64(stk_mask[sp]);
        } else {
            top_node = WALK_DONE;
            top_mask = 0;
        }
    };
    // The top of the parked stack is never below the cursor: a parked index
    // is the skip of an ancestor of the cursor's node, and no move inside
    // that subtree passes it.  So parked lanes rejoin exactly when the
    // cursor reaches their index, and the walk ends at total (where lanes
    // parked at total have rejoined first).  `active` only shrinks to the
    // lanes opening a node, which are never none, so a wave with a valid
    // lane keeps one.
    // The loop tests for the end at the bottom only, so the visitor's sums
    // stay in the same registers from one visit to the next.
    if(active == 0 || total <= 0) return wc;
    const char *gnb = reinterpret_cast<const char *>(gn);
    do {
        if constexpr(STATS) ++wc.visits;
        // the whole 64-byte record in one scalar load (the compiler would
        // otherwise split it by the fields used into five); the empty asm
        // statement only keeps all sixteen words live
        GravNode g;
        {
            typedef uint32_t u16v __attribute__((ext_vector_type(16)));
            const u16v raw = *(const __attribute__((address_space(4))) u16v *)(gnb + ((uint32_t)node << 6));
            asm volatile("" ::"s"(raw));
            __builtin_memcpy(&g, &raw, sizeof(GravNode));
        }
        const int skip = g.skip;
        const uint64_t om = v.node(g, active);
        if constexpr(STATS) {
            const int nopen = __popcll(om);
            wc.opened += nopen;
            wc.accepted += __popcll(active) - nopen;
        }
        if(om == 0) {
            node = skip;
        } else if(g.leaf) {
            v.leaf(g.begin, g.end, om);
            if constexpr(STATS) {
                wc.pairs += __popcll(om) * (g.end - g.begin);
                wc.slots += g.end - g.begin;
            }
            node = skip;
        } else {
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
        }
        while(node == top_node) {
            active |= top_mask;
            pop();
        }
    } while(node < total);
    return wc;
}

}  // namespace sphgpu
