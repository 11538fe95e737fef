// p2p_queue.hpp -- how a wave of the item kernels obtains its next work item.
//
// Persistent mode: the plan's items are split into n_queues queues (one per XCD by default).  A wave
// starts on queue blockIdx.x % n_queues -- workgroups are dispatched to the XCDs round-robin, so
// that is the queue of the wave's own XCD, whose items share sources in that XCD's L2 -- and takes
// items with one atomic per item.  When its queue is empty it moves to the next queue, and after
// it has found every queue empty it exits: at most n_queues moves, so every wave reaches the exit.
// Counters may overshoot their queue's length; the host resets them before each launch.
// Otherwise (counter null): one workgroup per item, item = blockIdx.x.
#pragma once

#include <hip/hip_runtime.h>

#include "p2p_internal.hpp"

namespace stkp2p {

struct ItemCursor {
    int q;
    int moves;
    bool first;
    __device__ explicit ItemCursor(const ItemArgs &a)
        : q(a.counter ? (int)(blockIdx.x % (unsigned)a.n_queues) : 0), moves(0), first(true) {}
};

// next item of this wave (uniform across the wave), or -1 when there is none left
__device__ inline int next_item(const ItemArgs &a, ItemCursor &c, int lane) {
    if (!a.counter) {
        if (!c.first)
            return -1;
        c.first = false;
        return (int)blockIdx.x < a.n_items ? (int)blockIdx.x : -1;
    }
    for (;;) {
        int v = 0;
        if (lane == 0)
            v = atomicAdd(a.counter + c.q, 1);
        v = __builtin_amdgcn_readfirstlane(v);
        const int it = a.q_ptr[c.q] + v;
        if (it < a.q_ptr[c.q + 1])
            return it;
        if (++c.moves >= a.n_queues)
            return -1;
        c.q = (c.q + 1 == a.n_queues) ? 0 : c.q + 1;
    }
}

} // namespace stkp2p
