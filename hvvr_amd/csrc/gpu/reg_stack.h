#pragma once
// Distance-sorted traversal stack held in one VGPR pair of a wave64 (the tile walk of
// traverse_avx.h:412-527, replayed over LDS records).
//
// Layout is bottom-up: stack position q (0 = bottom, size-1 = top, the closest entry)
// lives in lane q of (t, m); t = distance key, m = the entry's packed record meta.
// Capacity 64. A pop reads the top lane; an insert shifts the lanes above the insertion
// point up by one. Every operation is wave-uniform: all 64 lanes must be active.

#include "wave_traverse.h"

namespace hvvr {

enum : uint32_t { REG_STACK_CAP = 64 };

struct RegStack {
    float t;
    uint32_t m;
    uint32_t size;
};

__device__ inline uint32_t readLane(uint32_t v, uint32_t l) {
    return uint32_t(__builtin_amdgcn_readlane(int(v), int(l)));
}

// v with lane `l` replaced by the wave-uniform value `x`
__device__ inline uint32_t writeLane(uint32_t v, uint32_t x, uint32_t l, uint32_t lane) {
    return lane == l ? x : v;
}

// Whole-wave lane shift (DPP wave_shr:1): lane i takes lane i-1, lane 0 keeps `fill`. A DPP
// read from an inactive lane yields the old value, so the shift must run with all 64
// lanes active. Written as volatile asm, it stays where it is written (always
// wave-uniform code here); as a builtin under a select, the compiler may move it into a
// branch that masks lanes off. The s_nop gives the two wait states a DPP read needs
// after a VALU write of its source.
__device__ inline uint32_t laneShiftUp(uint32_t v, uint32_t fill) {
    uint32_t out = fill;
    asm volatile("s_nop 1\n\tv_mov_b32_dpp %0, %1 wave_shr:1 row_mask:0xf bank_mask:0xf" : "+v"(out) : "v"(v));
    return out;
}

// removes the top entry and returns its meta (size > 0)
__device__ inline uint32_t stackPop(RegStack& st) {
    st.size--;
    return readLane(st.m, st.size);
}

__device__ inline float stackTopKey(const RegStack& st) {
    return bitsf(readLane(fbits(st.t), st.size - 1u));
}

// Inserts (key, meta) one above the highest entry whose key is not smaller, so the new
// entry lands below the run of closer entries at the top (traverse_avx.h:505-515; the
// sentinel below the bottom is never smaller). The caller keeps size < 64.
__device__ inline void stackInsert(RegStack& st, float key, uint32_t meta, uint32_t lane) {
    const uint64_t b = ballot(lane < st.size && !(st.t < key));
    const uint32_t p = b ? 64u - uint32_t(__builtin_clzll(b)) : 0u;
    const float ut = bitsf(laneShiftUp(fbits(st.t), 0u));
    const uint32_t um = laneShiftUp(st.m, 0u);
    st.t = lane < p ? st.t : (lane == p ? key : ut);
    st.m = lane < p ? st.m : (lane == p ? meta : um);
    st.size++;
}

} // namespace hvvr
