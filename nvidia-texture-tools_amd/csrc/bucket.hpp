// bucket.hpp -- scratch layout and bookkeeping shared by the two-pass block
// encoders (BC3 colour, BC1 cluster fit): a front pass appends every block
// that needs a search to the list of its colour count, and the search pass
// walks the lists from the largest count down, so the lanes of a wave share
// their loop trip counts.
//
// Layout (32-bit words): counts[0, kCountSlots) per-count tallies;
// counts[kTopSlot] and kTicketGroups ticket counters kTicketStride words
// apart from kTicketBase (own cache lines, so their atomics do not queue
// behind one another); the lists start at kListsOffset, kCountSlots lists of
// `cap` block indices each.
#pragma once

#include <hip/hip_runtime.h>

namespace ntx_bucket {

constexpr int kCountSlots = 17;
constexpr int kTopSlot = 31;
constexpr int kTicketGroups = 16, kTicketBase = 32, kTicketStride = 32;
constexpr int kListsOffset = kTicketBase + kTicketGroups * kTicketStride;

// End of a search workgroup: the last one to finish clears the per-count
// tallies for the next image.  Every workgroup has used the tallies it
// loaded before it takes its ticket, so once all tickets are taken no one
// reads them again.  Two levels of tickets: workgroup b counts into group
// b % kTicketGroups, the last of a group into the top counter, which keeps
// the number of atomics on any one address small when thousands of
// workgroups end together.  No fence: nothing is published through the
// tickets, and the zeros reach the next kernel at the kernel boundary.
__device__ __forceinline__ void release_counts(int *counts)
{
    if (threadIdx.x == 0) {
        const int groups = min(int(gridDim.x), kTicketGroups);
        const int g = int(blockIdx.x) % kTicketGroups;
        const int members = (int(gridDim.x) - g + kTicketGroups - 1) / kTicketGroups;
        int *ticket = counts + kTicketBase + g * kTicketStride;
        if (atomicAdd(ticket, 1) == members - 1) {
            *ticket = 0;
            if (atomicAdd(&counts[kTopSlot], 1) == groups - 1) {
                for (int i = 0; i < kCountSlots; i++) counts[i] = 0;
                counts[kTopSlot] = 0;
            }
        }
    }
}

// Front pass, after every thread of the workgroup has its count k (and
// local = its position among the workgroup's blocks of count k, from an LDS
// tally s_cnt): one global atomic per count reserves the workgroup's run of
// each list.  Call with the whole workgroup; s_base is LDS scratch.
__device__ __forceinline__ void append_to_lists(int *s_cnt, int *s_base, int *counts, int *lists, int cap, int k,
                                                int local, int blk)
{
    const int t = threadIdx.x;
    __syncthreads();
    if (t < kCountSlots && s_cnt[t] > 0) s_base[t] = atomicAdd(&counts[t], s_cnt[t]);
    __syncthreads();
    if (local >= 0) lists[size_t(k) * cap + s_base[k] + local] = blk;
}

}  // namespace ntx_bucket
