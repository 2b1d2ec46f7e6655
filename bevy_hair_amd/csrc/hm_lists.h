// Device side of the covered-pixel sub-lists (PixelList, hm_kernels.h):
// writers append with one atomic per workgroup on the sub-list's counter,
// readers walk the concatenation through a prefix of the counts.
#pragma once

#include "hm_device.h"
#include "hm_kernels.h"

namespace hm {

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Appends each flagged thread's value to sub-list `sub` of L with one global
// atomic per workgroup. Every thread of the (256-thread) block calls it;
// `scratch` is 5 words of LDS. Returns nothing; entries past the capacity are
// dropped (the host sizes the sub-lists so that this cannot happen).
__device__ __forceinline__ void block_append(const PixelList& L, uint32_t sub, bool flag, uint32_t value,
                                             uint32_t* scratch) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(flag);
    if (lane == 0) scratch[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t n = scratch[0] + scratch[1] + scratch[2] + scratch[3];
        scratch[4] = n ? atomicAdd(L.counts + sub, n) : 0u;
    }
    __syncthreads();
    uint32_t pos = scratch[4];
    for (uint32_t w = 0; w < wave; ++w) pos += scratch[w];
    pos += lane_rank(m);
    if (flag && pos < L.cap) L.list[(size_t)sub * L.cap + pos] = value;
    __syncthreads();  // scratch is reused by the next call
}

// Exclusive prefix of the LIST_SUBS counts (clamped to cap) into pre[0..64];
// every thread calls it; pre[LIST_SUBS] is the total.
__device__ __forceinline__ void list_prefix(const PixelList& L, uint32_t* pre) {
    if (threadIdx.x < 64u) {
        uint32_t c = L.counts[threadIdx.x];
        c = c < L.cap ? c : L.cap;
        uint32_t inc = c;
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t y = __shfl_up(inc, d);
            if ((int)threadIdx.x >= d) inc += y;
        }
        pre[threadIdx.x + 1] = inc;
        if (threadIdx.x == 0) pre[0] = 0u;
    }
    __syncthreads();
}

// Entry i (< pre[LIST_SUBS]) of the concatenated sub-lists.
__device__ __forceinline__ uint32_t list_entry(const PixelList& L, const uint32_t* pre, uint32_t i) {
    uint32_t lo = 0u, hi = LIST_SUBS;  // last s with pre[s] <= i
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= i) lo = mid;
        else hi = mid;
    }
    return L.list[(size_t)lo * L.cap + (i - pre[lo])];
}

}  // namespace hm
